// host_batch.cpp -- populations: many independent 2D simulations stepped together, and
// their gradient tape.
//
// The reference evaluates a population by building one CpuPropagator per candidate and
// stepping each on its own thread (FitnessEvaluator::evaluate, fitness.rs:26-54, under
// rayon in EvolutionEngine::evaluate_population, search.rs:153-172).  Here all members live
// in one device arena and a step is a handful of launches: for every size class (same W, H)
// one launch per FFT pass walks a work-item list (member, block, slot); the fused stencil
// is one launch per channel count.  Both step parities are captured once into hipGraphs.
// A member whose flow leaves the fused stencil's radius is handled like hp_advance does:
// the whole batch is replayed from a snapshot with per-step checks and the general path.
//
// Gradient tape (no reference counterpart): tape_forward runs T steps keeping the state
// every k steps (and the final state); tape_backward walks the segments from the last,
// recomputes each segment's states from its checkpoint, and runs the adjoint of every step
// in reverse -- the forward's three FFT passes (keeping the convolutions), the flow and
// divergence adjoints, and the three adjoint FFT passes -- accumulating dL/dA, the growth
// parameter sums and the kernel-tap spectrum.  The ring parameters' derivatives follow on
// the host in FP64 from the kernel-tap derivatives.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <exception>
#include <map>
#include <mutex>
#include <string>
#include <thread>
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
    std::vector<int> radius;  // actual kernel radius per kernel
    float time = 0.0f;
    uint64_t step = 0;
};

struct ItemList {
    int4* d = nullptr;
    int n = 0;
};

struct SizeClass {
    int W, H;
    ItemList rows_fwd, spectral, rows_inv, adj_rows_fwd, adj_spectral;
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
    ItemList copy_items;  // (member, part, nparts) over each member's C*N floats
    ItemList adj_tiles;   // (member, tile) of the adjoint stencil passes
    ItemList wdft_rows;   // (member, kernel, row block, partial offset) of the tape's layout
    ItemList wdft_sums;   // (member, kernel, row blocks, partial offset)
    double* d_wpart = nullptr;  // the kernel-tap windows of every row block
    std::map<int, float2*> tw;
    std::vector<void*> allocs;
    std::vector<void*> tape_allocs;
    int* d_flags = nullptr;  // [0] stencil flags, [1] needed radius, [2] non-finite
    float4* box_lo = nullptr;
    float4* box_hi = nullptr;
    double* d_moments = nullptr;
    double* d_partials = nullptr;
    double* d_mass = nullptr;
    int cur = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    // step mode: one launch per pass and class (default) or one persistent launch per step
    bool persistent = false;
    bool persistent_ok = false;  // every member's sides are compiled lengths of the persistent step
    ItemList pitems;             // the persistent step's item segments, back to back
    int pstart[32] = {0};        // segment starts (batch_step_persistent_segments() + 1 used)
    unsigned* d_bar = nullptr;   // its grid-barrier counters (zero between launches)
    uint64_t barrier_timeouts = 0;
    uint32_t max_radius = 0;
    uint64_t general_steps = 0;
    // plan: LDS FFT variant and stencil tile shared by every member
    hp::FftVariant fv;
    int tile = 0;
    std::map<std::pair<int, int>, std::vector<uint32_t>> by_size;      // (W, H) -> members
    std::map<std::pair<int, bool>, std::vector<uint32_t>> by_stencil;  // (C, pow2 area) -> members
    std::vector<void*> item_allocs;  // work-item lists of the current plan
    // tape
    float* d_stage = nullptr;     // contiguous device staging of every member's C*N values
    float* h_stage = nullptr;     // pinned host twin of d_stage
    double* d_pacc = nullptr;     // every member's growth-parameter sums, back to back
    double* d_dkwin = nullptr;    // every member's kernel-tap windows, back to back
    size_t stage_n = 0, pacc_n = 0, dkwin_n = 0;
    uint64_t tape_cap_slots = 0;  // states each member's tape holds
    uint64_t conv_cap_slots = 0;  // convolution slots (K*N each) each member holds
    bool tape_conv = false;       // the tape keeps every step's convolutions
    uint64_t tape_T = 0;
    uint32_t tape_k = 0;
    bool tape_recorded = false;
    int tape_cur = 0;  // parity of the final state after tape_forward
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

#define STRY(expr)                      \
    do {                                \
        hp_status s_ = (expr);          \
        if (s_ != HP_OK) return s_;     \
    } while (0)

namespace {

template <typename T>
hp_status balloc(hp_batch* b, T** p, size_t count, std::vector<void*>* list = nullptr) {
    void* v = nullptr;
    BTRY(b, hipMalloc(&v, std::max<size_t>(1, count) * sizeof(T)));
    (list ? list : &b->allocs)->push_back(v);
    *p = (T*)v;
    return HP_OK;
}

template <typename T>
hp_status bupload(hp_batch* b, T** p, const std::vector<T>& v, std::vector<void*>* list = nullptr) {
    hp_status st = balloc(b, p, v.size(), list);
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

// The three FFT passes of a step from A[cur]; conv: also store the convolutions in conv slot
// cslot; rows_only: the forward row pass alone (the adjoint's source spectra).
void enqueue_fft_passes(hp_batch* b, int cur, bool conv, int cslot = 0, bool rows_only = false) {
    for (const SizeClass& sc : b->classes) {
        if (sc.rows_fwd.n)
            hp::launch_batch_rows_fwd(sc.W, b->stream, b->d_desc, sc.rows_fwd.d, sc.rows_fwd.n, cur, b->fv);
        if (rows_only) continue;
        if (sc.spectral.n) hp::launch_batch_spectral(sc.H, b->stream, b->d_desc, sc.spectral.d, sc.spectral.n, b->fv);
        if (sc.rows_inv.n)
            hp::launch_batch_rows_inv(sc.W, b->stream, b->d_desc, sc.rows_inv.d, sc.rows_inv.n, conv, cslot, b->fv);
    }
}

void enqueue_stencils(hp_batch* b, int cur) {
    for (const StencilGroup& gr : b->stencil_groups)
        hp::launch_batch_stencil2d(b->stream, gr.C, gr.pow2, b->d_desc, gr.items.d, gr.items.n, cur, b->d_flags,
                                   b->tile);
}

// Bound on a grid barrier's polls (a few tenths of a second of sleeps): far past any phase's
// length, short enough that a grid that cannot be resident ends the launch quickly.
constexpr int kBarrierSpins = 1 << 20;

hp_status enqueue_step(hp_batch* b, int cur) {
    if (b->persistent) {
        hp::launch_batch_step_persistent(b->stream, b->d_desc, b->pitems.d, b->pstart, cur, b->d_flags, b->d_bar,
                                         std::min(hp::batch_step_persistent_grid(), std::max(1, b->pitems.n)),
                                         kBarrierSpins);
    } else {
        enqueue_fft_passes(b, cur, false);
        enqueue_stencils(b, cur);
    }
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
// U: the affinity to read (null: the member's own U buffer).
hp_status member_stencil(hp_batch* b, int i, int cur, const float* U = nullptr) {
    const hp::MemberDesc& m = b->desc[i];
    const HostMember& hm = b->members[i];
    const float* u = U ? U : m.U;
    const float s = hm.cfg.flow.distribution_size;
    bool general = !hp::stencil2d_supported(m.g, s, b->tile);
    if (!general) {
        hp::launch_clear_ints(b->stream, b->d_flags, 1);
        hp::launch_stencil2d(b->stream, m.A[cur], u, m.A[1 - cur], m.C, m.g, m.dt, m.beta_a, m.nexp, s, b->d_flags,
                             b->tile);
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
    hp::launch_boxes(b->stream, m.A[cur], u, m.C, m.g, m.dt, m.beta_a, m.nexp, s, b->box_lo, b->box_hi,
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

// One step with every member's stencil checked (A[cur] -> A[1 - cur]).  u_slot >= 0: the device
// descriptors point U at tape_u slot u_slot (a taped forward), where the FFT passes leave it, so
// the host-launched stencils read it there too.
hp_status step_checked(hp_batch* b, int cur, bool conv = false, int cslot = 0, int u_slot = -1) {
    enqueue_fft_passes(b, cur, conv, cslot);
    for (size_t i = 0; i < b->members.size(); i++) {
        const hp::MemberDesc& m = b->desc[i];
        const float* u = u_slot >= 0 ? m.tape_u + (size_t)u_slot * m.C * m.g.N : nullptr;
        STRY(member_stencil(b, (int)i, cur, u));
    }
    return HP_OK;
}

void copy_states(hp_batch* b, int cur, int slot, bool to_tape) {
    hp::launch_tape_copy(b->stream, b->d_desc, b->copy_items.d, b->copy_items.n, cur, slot, to_tape);
}

hp_status read_flag(hp_batch* b, int* f) {
    BTRY(b, hipMemcpyAsync(f, b->d_flags, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    BTRY(b, hipStreamSynchronize(b->stream));
    return HP_OK;
}

hp_status advance(hp_batch* b, uint64_t steps) {
    if (steps == 0) return HP_OK;
    const int start = b->cur;
    copy_states(b, start, -1, true);  // snapshot
    hp::launch_clear_ints(b->stream, b->d_flags, 1);
    STRY(build_graphs(b));
    for (uint64_t k = 0; k < steps; k++) {
        BTRY(b, hipGraphLaunch(b->graph[b->cur], b->stream));
        b->cur = 1 - b->cur;
    }
    int f = 0;
    STRY(read_flag(b, &f));
    if (f & 4) b->barrier_timeouts++;  // a persistent step's grid barrier timed out: recompute
    if (f != 0) {  // replay from the snapshot, one checked step at a time
        b->cur = start;
        copy_states(b, start, -1, false);
        for (uint64_t k = 0; k < steps; k++) {
            STRY(step_checked(b, b->cur));
            b->cur = 1 - b->cur;
        }
    } else {
        b->max_radius = std::max<uint32_t>(b->max_radius, 2);
    }
    for (HostMember& hm : b->members)
        for (uint64_t k = 0; k < steps; k++) {
            hm.time += hm.cfg.dt;
            hm.step += 1;
        }
    return HP_OK;
}

// ---- gradient tape --------------------------------------------------------------------------
// Tape slots per member: [0, nck) checkpoints at t = j*k, nck the final state, then k
// segment states.
uint64_t num_checkpoints(uint64_t T, uint32_t k) { return (T + k - 1) / k; }

hp_status tape_check_member(hp_batch* b, size_t i) {
    const HostMember& hm = b->members[i];
    const hp::MemberPlan& p = hm.plan;
    const std::string who = "member " + std::to_string(i) + ": ";
    if (!hp::adj_fft_supported(p.g.W) || !hp::adj_fft_supported(p.g.H) || p.g.W > hp::window_dft_max_side() ||
        p.g.H > hp::window_dft_max_side())
        return b->fail(HP_ERR_UNSUPPORTED_SIZE, who + "gradient tapes need sides 2^k or 3 * 2^k up to " +
                                                   std::to_string(hp::window_dft_max_side()));
    if (p.spec_work.size() != p.src_chan.size())
        return b->fail(HP_ERR_UNSUPPORTED_SIZE, who + "gradient tapes need one spectral item per source");
    for (int r : hm.radius)
        if (r < 1 || 2 * r + 1 > std::min(p.g.W, p.g.H) || 2 * r + 1 > hp::window_dft_max_taps())
            return b->fail(HP_ERR_UNSUPPORTED_SIZE,
                           who + "gradient tapes need every kernel's support (2R+1) within the grid and <= " +
                               std::to_string(hp::window_dft_max_taps()));
    return HP_OK;
}

hp_status ensure_tape(hp_batch* b, uint64_t T, uint32_t k, bool store_conv) {
    const uint64_t slots = num_checkpoints(T, k) + 1 + k;
    const uint64_t cslots = store_conv ? std::max<uint64_t>(T, 1) : 1;
    if (slots <= b->tape_cap_slots && cslots <= b->conv_cap_slots) return HP_OK;
    for (size_t i = 0; i < b->members.size(); i++) STRY(tape_check_member(b, i));
    BTRY(b, hipStreamSynchronize(b->stream));
    for (void* p : b->tape_allocs) (void)hipFree(p);
    b->tape_allocs.clear();
    b->tape_cap_slots = 0;
    for (size_t i = 0; i < b->members.size(); i++) {
        const HostMember& hm = b->members[i];
        const hp::MemberPlan& p = hm.plan;
        hp::MemberDesc& d = b->desc[i];
        const size_t N = (size_t)p.g.N, CN = (size_t)p.C * N, K = (size_t)std::max(p.K, 0);
        std::vector<void*>* L = &b->tape_allocs;
        STRY(balloc(b, &d.lam[0], CN, L));
        STRY(balloc(b, &d.lam[1], CN, L));
        STRY(balloc(b, &d.lamU, CN, L));
        STRY(balloc(b, &d.conv, cslots * K * N, L));
        STRY(balloc(b, &d.tape_u, cslots * CN, L));  // as many slots as conv: conv_cap_slots covers these
        STRY(balloc(b, &d.tape_x, cslots * std::max<size_t>(1, p.src_chan.size()) * (size_t)p.g.spec, L));
        STRY(balloc(b, &d.gscr, (2 * (size_t)p.C + 3) * N, L));
        STRY(balloc(b, &d.P, K * (size_t)p.g.spec, L));
        STRY(balloc(b, &d.tape, slots * CN, L));
        std::vector<int2> km(std::max<size_t>(K, 1)), dkw(std::max<size_t>(K, 1));
        size_t off = 0;
        for (size_t kk = 0; kk < K; kk++) {
            km[kk] = make_int2((int)hm.kernels[kk].source_channel, (int)hm.kernels[kk].target_channel);
            const int R = hm.radius[kk];
            dkw[kk] = make_int2((int)off, R);
            off += (size_t)(2 * R + 1) * (2 * R + 1);
        }
        int2* dkm = nullptr;
        int2* ddk = nullptr;
        STRY(bupload(b, &dkm, km, L));
        STRY(bupload(b, &ddk, dkw, L));
        d.kmap = dkm;
        d.dkw = ddk;
    }
    // contiguous accumulators and transfer staging, so that results come back in one copy each
    size_t pacc_n = 0, dkwin_n = 0, stage_n = 0;
    for (size_t i = 0; i < b->members.size(); i++) {
        const HostMember& hm = b->members[i];
        const hp::MemberPlan& p = hm.plan;
        pacc_n += (size_t)std::max(p.K, 0) * ((p.g.H + 1) / 2) * 3;
        for (int r : hm.radius) dkwin_n += (size_t)(2 * r + 1) * (2 * r + 1);
        stage_n += (size_t)p.C * p.g.N;
    }
    // the window DFT's row-block items and partial windows
    {
        std::vector<int4> rows, sums;
        size_t poff = 0;
        const int RB = hp::window_dft_rows();
        for (size_t i = 0; i < b->members.size(); i++) {
            const HostMember& hm = b->members[i];
            const int nrb = (hm.plan.g.H + RB - 1) / RB;
            for (size_t kk = 0; kk < hm.radius.size(); kk++) {
                const size_t SS = (size_t)(2 * hm.radius[kk] + 1) * (2 * hm.radius[kk] + 1);
                if (poff + nrb * SS > (size_t)0x7fffffff)
                    return b->fail(HP_ERR_UNSUPPORTED_SIZE, "kernel-tap windows beyond 2^31 values");
                for (int rb = 0; rb < nrb; rb++) rows.push_back(make_int4((int)i, (int)kk, rb, (int)poff));
                sums.push_back(make_int4((int)i, (int)kk, nrb, (int)poff));
                poff += nrb * SS;
            }
        }
        STRY(bupload(b, &b->wdft_rows.d, rows, &b->tape_allocs));
        b->wdft_rows.n = (int)rows.size();
        STRY(bupload(b, &b->wdft_sums.d, sums, &b->tape_allocs));
        b->wdft_sums.n = (int)sums.size();
        STRY(balloc(b, &b->d_wpart, std::max<size_t>(1, poff), &b->tape_allocs));
    }
    STRY(balloc(b, &b->d_pacc, pacc_n, &b->tape_allocs));
    STRY(balloc(b, &b->d_dkwin, dkwin_n, &b->tape_allocs));
    STRY(balloc(b, &b->d_stage, stage_n, &b->tape_allocs));
    if (b->h_stage) (void)hipHostFree(b->h_stage);
    b->h_stage = nullptr;
    BTRY(b, hipHostMalloc((void**)&b->h_stage, std::max<size_t>(1, stage_n) * sizeof(float), hipHostMallocDefault));
    b->pacc_n = pacc_n;
    b->dkwin_n = dkwin_n;
    b->stage_n = stage_n;
    pacc_n = dkwin_n = stage_n = 0;
    for (size_t i = 0; i < b->members.size(); i++) {
        const HostMember& hm = b->members[i];
        const hp::MemberPlan& p = hm.plan;
        hp::MemberDesc& d = b->desc[i];
        d.pacc = b->d_pacc + pacc_n;
        d.dkwin = b->d_dkwin + dkwin_n;
        d.stage_off = (long long)stage_n;
        pacc_n += (size_t)std::max(p.K, 0) * ((p.g.H + 1) / 2) * 3;
        for (int r : hm.radius) dkwin_n += (size_t)(2 * r + 1) * (2 * r + 1);
        stage_n += (size_t)p.C * p.g.N;
    }
    BTRY(b, hipMemcpy(b->d_desc, b->desc.data(), b->desc.size() * sizeof(hp::MemberDesc), hipMemcpyHostToDevice));
    b->tape_cap_slots = slots;
    b->conv_cap_slots = cslots;
    return HP_OK;
}

// Tape layout for checkpoint_every = 0, from the free device memory (half of it is the
// budget): every state and every step's convolutions if they fit (no recomputation at all,
// and the backward skips the forward's spectral passes); else every state; else a
// checkpoint every ~sqrt(T) steps (about 2 sqrt(T) states per member).
void auto_layout(hp_batch* b, uint64_t T, uint32_t* k, bool* store_conv) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    double set = 0.0, cset = 0.0;  // bytes of one state / one step's convolutions and U, all members
    for (const HostMember& hm : b->members) {
        set += 4.0 * hm.plan.C * (double)hm.plan.g.N;
        cset += 4.0 * (std::max(hm.plan.K, 0) + hm.plan.C) * (double)hm.plan.g.N +
                8.0 * (double)hm.plan.src_chan.size() * (double)hm.plan.g.spec;
    }
    const double budget = 0.5 * (double)free_b;
    *store_conv = false;
    if ((double)(T + 3) * (set + cset) <= budget) {
        *k = 1;
        *store_conv = true;
    } else if ((double)(T + 3) * set <= budget) {
        *k = 1;
    } else {
        *k = (uint32_t)std::max<uint64_t>(1, (uint64_t)ceil(sqrt((double)std::max<uint64_t>(T, 1))));
    }
}

// T steps from the current state, keeping a state every k steps (and the final one); with
// store_conv every step's convolutions too.  The steps run back to back; if any member
// left the fused stencil's radius the whole tape is redone with checked steps.
hp_status tape_forward(hp_batch* b, uint64_t T, uint32_t k) {
    bool store_conv = false;
    if (k == 0) auto_layout(b, T, &k, &store_conv);
    b->tape_recorded = false;
    STRY(ensure_tape(b, T, k, store_conv));
    const uint64_t nck = num_checkpoints(T, k);
    const int start = b->cur;
    if (!store_conv) STRY(build_graphs(b));
    for (int pass = 0; pass < 2; pass++) {
        int cur = start;
        if (pass == 1) copy_states(b, start, 0, false);  // back to the first checkpoint
        hp::launch_clear_ints(b->stream, b->d_flags, 1);
        hp_status lst = HP_OK;
        for (uint64_t t = 0; t < T && lst == HP_OK; t++) {
            if (t % k == 0) copy_states(b, cur, (int)(t / k), true);
            // with stored convolutions the step's U and source spectra go straight into their
            // tape slots (no graph runs in this loop; the descriptors are restored after it)
            if (store_conv) hp::launch_point_slots(b->stream, b->d_desc, (int)b->members.size(), -1, (int)t);
            if (pass == 1) {
                lst = step_checked(b, cur, store_conv, (int)t, store_conv ? (int)t : -1);
            } else if (store_conv) {
                enqueue_fft_passes(b, cur, true, (int)t);
                enqueue_stencils(b, cur);
            } else if (hipGraphLaunch(b->graph[cur], b->stream) != hipSuccess) {
                lst = b->fail(HP_ERR_HIP, "hipGraphLaunch failed");
            }
            cur = 1 - cur;
        }
        if (store_conv)
            BTRY(b, hipMemcpyAsync(b->d_desc, b->desc.data(), b->desc.size() * sizeof(hp::MemberDesc),
                                   hipMemcpyHostToDevice, b->stream));
        if (lst != HP_OK) {  // the population goes back to the state the call started from
            copy_states(b, start, 0, false);
            b->cur = start;
            (void)hipStreamSynchronize(b->stream);
            return lst;
        }
        BTRY(b, hipGetLastError());
        copy_states(b, cur, (int)nck, true);
        b->cur = cur;
        int f = 0;
        STRY(read_flag(b, &f));
        if (pass == 0 && f == 0) {
            if (T) b->max_radius = std::max<uint32_t>(b->max_radius, 2);
            break;
        }
    }
    for (HostMember& hm : b->members)
        for (uint64_t t = 0; t < T; t++) {
            hm.time += hm.cfg.dt;
            hm.step += 1;
        }
    b->tape_T = T;
    b->tape_k = k;
    b->tape_conv = store_conv;
    b->tape_cur = b->cur;
    b->tape_recorded = true;
    return HP_OK;
}

// States t0+1 .. t1-1 into segment slots seg + (t - t0), from checkpoint j (the state at
// t0 = j*k); A[0]/A[1] are scratch here.  Nothing to do for one-step segments.
hp_status recompute_segment(hp_batch* b, uint64_t j, uint64_t t0, uint64_t t1) {
    if (t1 - t0 <= 1) return HP_OK;
    const int seg = (int)(num_checkpoints(b->tape_T, b->tape_k) + 1);
    STRY(build_graphs(b));
    for (int pass = 0; pass < 2; pass++) {
        int cur = 0;
        copy_states(b, 0, (int)j, false);
        hp::launch_clear_ints(b->stream, b->d_flags, 1);
        for (uint64_t t = t0 + 1; t < t1; t++) {
            if (pass == 0) BTRY(b, hipGraphLaunch(b->graph[cur], b->stream));
            else STRY(step_checked(b, cur));
            cur = 1 - cur;
            copy_states(b, cur, seg + (int)(t - t0), true);
        }
        int f = 0;
        STRY(read_flag(b, &f));
        if (pass == 0 && f == 0) break;  // the fused stencil covered every step
        if (pass == 1) break;
    }
    return HP_OK;
}

// Adjoint of step t whose input state is in A[0]; lam[li] holds dL/dA_{t+1}, lam[1-li]
// receives dL/dA_t.  With stored convolutions only the forward row pass (the source spectra)
// and U are rebuilt; otherwise the step's FFT passes run again, keeping the convolutions.
// u_in_place: the descriptors already point U at the step's tape slot.
hp_status backward_step(hp_batch* b, int li, uint64_t t, bool u_in_place = false) {
    int cslot = 0;
    if (b->tape_conv) {
        cslot = (int)t;
        if (!u_in_place) {  // otherwise the descriptors point X and U at the step's slots
            enqueue_fft_passes(b, 0, false, 0, true);
            hp::launch_u_copy(b->stream, b->d_desc, b->copy_items.d, b->copy_items.n, cslot, false);
        }
    } else {
        enqueue_fft_passes(b, 0, true, 0);
    }
    hp::launch_adj_flow(b->stream, b->d_desc, b->adj_tiles.d, b->adj_tiles.n, 0, li);
    hp::launch_adj_div(b->stream, b->d_desc, b->adj_tiles.d, b->adj_tiles.n, li);
    for (const SizeClass& sc : b->classes) {
        if (sc.adj_rows_fwd.n)
            hp::launch_adj_rows_fwd(sc.W, b->stream, b->d_desc, sc.adj_rows_fwd.d, sc.adj_rows_fwd.n, cslot);
        if (sc.adj_spectral.n)
            hp::launch_adj_spectral(sc.H, b->stream, b->d_desc, sc.adj_spectral.d, sc.adj_spectral.n);
        if (sc.rows_fwd.n) hp::launch_adj_rows_inv(sc.W, b->stream, b->d_desc, sc.rows_fwd.d, sc.rows_fwd.n, li);
    }
    BTRY(b, hipGetLastError());
    return HP_OK;
}

// theta = per kernel (weight, mu, sigma), then per ring (amplitude, distance, width).
size_t theta_count(const HostMember& hm) { return 3 * (size_t)hm.cfg.num_kernels + 3 * (size_t)hm.cfg.num_rings; }

// dL/dtheta of one member from the device sums: growth sums reduced over row pairs; ring
// parameters through K = v / sum(v), v = sum_r a_r exp(-(rho - d_r)^2 / (2 w_r^2)) on
// rho = |offset| / R <= 1 (kernel.rs:40-72), in FP64.
void member_theta(const HostMember& hm, const std::vector<double>& pacc, const std::vector<double>& dkwin,
                  double* th) {
    const int K = (int)hm.cfg.num_kernels;
    const int pairs = (hm.plan.g.H + 1) / 2;
    std::fill(th, th + theta_count(hm), 0.0);
    for (int k = 0; k < K; k++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int p = 0; p < pairs; p++) s += pacc[((size_t)k * pairs + p) * 3 + j];
            th[3 * k + j] = s;
        }
    const double invN = 1.0 / (double)hm.plan.g.N;
    size_t off = 0;
    for (int k = 0; k < K; k++) {
        const hp_kernel& kc = hm.kernels[k];
        const int R = hm.radius[k], S = 2 * R + 1;
        const double* dk = dkwin.data() + off;
        off += (size_t)S * S;
        std::vector<double> v((size_t)S * S, 0.0);
        std::vector<char> in((size_t)S * S, 0);
        const size_t nr = kc.ring_count;
        std::vector<double> ex((size_t)S * S * nr, 0.0);  // each tap's ring exponentials, reused below
        double sum = 0.0, dot = 0.0;
        for (int y = 0; y < S; y++)
            for (int x = 0; x < S; x++) {
                const double dx = x - R, dy = y - R;
                const double rho = sqrt(dx * dx + dy * dy) / R;
                if (rho > 1.0) continue;
                double val = 0.0;
                for (uint32_t r = 0; r < kc.ring_count; r++) {
                    const hp_ring& rg = hm.rings[kc.ring_offset + r];
                    const double rd = rho - rg.distance, w = rg.width;
                    const double e = exp(-rd * rd / (2.0 * w * w));
                    ex[((size_t)y * S + x) * nr + r] = e;
                    val += rg.amplitude * e;
                }
                v[(size_t)y * S + x] = val;
                in[(size_t)y * S + x] = 1;
                sum += val;
                dot += dk[(size_t)y * S + x] * invN * val;
            }
        for (int y = 0; y < S; y++)
            for (int x = 0; x < S; x++) {
                if (!in[(size_t)y * S + x]) continue;
                const double g = dk[(size_t)y * S + x] * invN;
                const double dv = sum > 0.0 ? g / sum - dot / (sum * sum) : g;
                const double dx = x - R, dy = y - R;
                const double rho = sqrt(dx * dx + dy * dy) / R;
                for (uint32_t r = 0; r < kc.ring_count; r++) {
                    const hp_ring& rg = hm.rings[kc.ring_offset + r];
                    const double a = rg.amplitude, rd = rho - rg.distance, w = rg.width;
                    const double e = ex[((size_t)y * S + x) * nr + r];
                    double* tr = th + 3 * (size_t)K + 3 * (size_t)(kc.ring_offset + r);
                    tr[0] += dv * e;
                    tr[1] += dv * a * e * rd / (w * w);
                    tr[2] += dv * a * e * rd * rd / (w * w * w);
                }
            }
    }
}

// f(i) for every member i over up to 16 host threads (callers give each member its own outputs).
template <class F>
void parallel_members(size_t M, const F& f) {
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t nth = std::min<size_t>(M, std::max(1u, std::min(16u, hw)));
    if (nth <= 1) {
        for (size_t i = 0; i < M; i++) f(i);
        return;
    }
    // An exception in a worker (an allocation in f) is kept and rethrown on the calling thread
    // after the join, where the entry point's guard turns it into an error status; strides whose
    // thread could not be started run on the calling thread.
    std::mutex mu;
    std::exception_ptr first;
    auto stride = [&](size_t t) {
        try {
            for (size_t i = t; i < M; i += nth) f(i);
        } catch (...) {
            std::lock_guard<std::mutex> lk(mu);
            if (!first) first = std::current_exception();
        }
    };
    std::vector<std::thread> pool;
    size_t started = 0;
    try {
        for (; started < nth; started++) pool.emplace_back(stride, started);
    } catch (...) {
    }
    for (size_t t = started; t < nth; t++) stride(t);
    for (std::thread& th : pool) th.join();
    if (first) std::rethrow_exception(first);
}

hp_status tape_backward(hp_batch* b, const float* const* seed, float* const* gA0, double* dtheta) {
    if (!b->tape_recorded) return b->fail(HP_ERR_SHAPE, "no recorded tape: call hp_batch_tape_forward first");
    const uint64_t T = b->tape_T, k = b->tape_k, nck = num_checkpoints(T, b->tape_k);
    size_t ci = 0;
    std::vector<size_t> cfirst(b->members.size());
    for (size_t i = 0; i < b->members.size(); i++) {
        cfirst[i] = ci;
        for (int c = 0; c < b->desc[i].C; c++, ci++)
            if (!seed[ci]) return b->fail(HP_ERR_NULL, "dL/dA_T pointer is NULL");
    }
    parallel_members(b->members.size(), [&](size_t i) {
        const hp::MemberDesc& m = b->desc[i];
        for (int c = 0; c < m.C; c++)
            memcpy(b->h_stage + m.stage_off + (size_t)c * m.g.N, seed[cfirst[i] + c], (size_t)m.g.N * sizeof(float));
    });
    BTRY(b, hipMemcpyAsync(b->d_stage, b->h_stage, b->stage_n * sizeof(float), hipMemcpyHostToDevice, b->stream));
    hp::launch_stage_lam(b->stream, b->d_desc, b->copy_items.d, b->copy_items.n, b->d_stage, 0, false);
    hp::launch_tape_clear(b->stream, b->d_desc, b->copy_items.d, b->copy_items.n);
    // With a checkpoint every step and stored convolutions nothing is recomputed and no graph
    // runs in the loop: each backward step reads its state and U straight from their tape slots
    // (the descriptors point there) instead of copying them into A[0] and U.  The descriptors
    // are restored right after the loop.
    const bool direct = k == 1 && b->tape_conv;
    int li = 0;
    hp_status lst = HP_OK;
    for (uint64_t j = nck; j-- > 0 && lst == HP_OK;) {
        const uint64_t t0 = j * k, t1 = std::min<uint64_t>(T, t0 + k);
        lst = recompute_segment(b, j, t0, t1);
        const int seg = (int)(nck + 1);
        for (uint64_t t = t1; t-- > t0 && lst == HP_OK;) {
            const int slot = t == t0 ? (int)j : seg + (int)(t - t0);
            if (direct) hp::launch_point_slots(b->stream, b->d_desc, (int)b->members.size(), slot, (int)t);
            else copy_states(b, 0, slot, false);
            lst = backward_step(b, li, t, direct);
            li = 1 - li;
        }
    }
    if (direct)  // also after a failed step: later calls must see the members' own buffers
        BTRY(b, hipMemcpyAsync(b->d_desc, b->desc.data(), b->desc.size() * sizeof(hp::MemberDesc),
                               hipMemcpyHostToDevice, b->stream));
    if (lst != HP_OK) {  // A[0]/A[1] served as scratch: the population holds A_T again
        copy_states(b, b->tape_cur, (int)nck, false);
        b->cur = b->tape_cur;
        (void)hipStreamSynchronize(b->stream);
        return lst;
    }
    hp::launch_window_dft(b->stream, b->d_desc, b->wdft_rows.d, b->wdft_rows.n, b->wdft_sums.d, b->wdft_sums.n,
                          b->d_wpart);
    BTRY(b, hipGetLastError());
    // results: one copy each for dL/dA_0, the growth sums and the kernel-tap windows
    hp::launch_stage_lam(b->stream, b->d_desc, b->copy_items.d, b->copy_items.n, b->d_stage, li, true);
    BTRY(b, hipGetLastError());
    BTRY(b, hipMemcpyAsync(b->h_stage, b->d_stage, b->stage_n * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    std::vector<double> pacc(std::max<size_t>(1, b->pacc_n)), dkwin(std::max<size_t>(1, b->dkwin_n));
    if (b->pacc_n)
        BTRY(b, hipMemcpyAsync(pacc.data(), b->d_pacc, b->pacc_n * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    if (b->dkwin_n)
        BTRY(b, hipMemcpyAsync(dkwin.data(), b->d_dkwin, b->dkwin_n * sizeof(double), hipMemcpyDeviceToHost,
                               b->stream));
    copy_states(b, b->tape_cur, (int)nck, false);  // the population holds A_T again
    b->cur = b->tape_cur;
    BTRY(b, hipStreamSynchronize(b->stream));
    // Every member's outputs are its own ranges of gA0 and dtheta, so the members are spread
    // over up to 16 host threads (the FP64 ring chain rule is most of the host's work here).
    const size_t M = b->members.size();
    std::vector<size_t> cbase(M), off(M), poff(M), woff(M);
    for (size_t i = 0, c0 = 0, o = 0, po = 0, wo = 0; i < M; i++) {
        const HostMember& hm = b->members[i];
        cbase[i] = c0;
        off[i] = o;
        poff[i] = po;
        woff[i] = wo;
        c0 += b->desc[i].C;
        o += theta_count(hm);
        po += (size_t)hm.cfg.num_kernels * ((b->desc[i].g.H + 1) / 2) * 3;
        for (int r : hm.radius) wo += (size_t)(2 * r + 1) * (2 * r + 1);
    }
    auto member_out = [&](size_t i) {
        const hp::MemberDesc& m = b->desc[i];
        const HostMember& hm = b->members[i];
        for (int c = 0; c < m.C; c++)
            if (gA0 && gA0[cbase[i] + c])
                memcpy(gA0[cbase[i] + c], b->h_stage + m.stage_off + (size_t)c * m.g.N, (size_t)m.g.N * sizeof(float));
        if (dtheta) {
            const size_t pn = (size_t)hm.cfg.num_kernels * ((m.g.H + 1) / 2) * 3;
            size_t wn = 0;
            for (int r : hm.radius) wn += (size_t)(2 * r + 1) * (2 * r + 1);
            std::vector<double> pm(pacc.begin() + poff[i], pacc.begin() + poff[i] + pn);
            std::vector<double> wm(dkwin.begin() + woff[i], dkwin.begin() + woff[i] + wn);
            member_theta(hm, pm, wm, dtheta + off[i]);
        }
    };
    parallel_members(M, member_out);
    return HP_OK;
}

// Work-item lists of the current plan: per size class the FFT passes' items (the spectral
// pass's column blocks depend on the plan's column split) and per stencil group the tiles of
// the plan's tile shape.
hp_status build_items(hp_batch* b) {
    // the captured graphs hold the old lists as kernel arguments: they go first, so that no
    // early return below can leave a graph pointing at freed lists
    BTRY(b, hipStreamSynchronize(b->stream));
    for (int p = 0; p < 2; p++)
        if (b->graph[p]) {
            (void)hipGraphExecDestroy(b->graph[p]);
            b->graph[p] = nullptr;
        }
    for (void* p : b->item_allocs) (void)hipFree(p);
    b->item_allocs.clear();
    b->classes.clear();
    b->stencil_groups.clear();
    std::vector<void*>* L = &b->item_allocs;
    for (auto& kv : b->by_size) {
        SizeClass scl{kv.first.first, kv.first.second, {}, {}, {}, {}, {}};
        std::vector<int4> rf, sp, ri, arf, asp;
        const int CB = hp::spectral_block_lines(scl.H, b->fv);
        const int ACB = std::max(1, hp::adj_spectral_block_lines(scl.H));
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
                for (int kk = 0; kk < p.K; kk++)
                    for (int pr = 0; pr < pairs; pr++) arf.push_back(make_int4((int)i, pr, kk, 0));
                for (int wi = 0; wi < (int)p.spec_work.size(); wi++)
                    for (int blk = 0; blk < (p.g.Wh + ACB - 1) / ACB; blk++)
                        asp.push_back(make_int4((int)i, blk, wi, 0));
            }
        }
        STRY(bupload(b, &scl.rows_fwd.d, rf, L));
        scl.rows_fwd.n = (int)rf.size();
        STRY(bupload(b, &scl.spectral.d, sp, L));
        scl.spectral.n = (int)sp.size();
        STRY(bupload(b, &scl.rows_inv.d, ri, L));
        scl.rows_inv.n = (int)ri.size();
        STRY(bupload(b, &scl.adj_rows_fwd.d, arf, L));
        scl.adj_rows_fwd.n = (int)arf.size();
        STRY(bupload(b, &scl.adj_spectral.d, asp, L));
        scl.adj_spectral.n = (int)asp.size();
        b->classes.push_back(scl);
    }
    const int TX = hp::stencil2d_tile_w(b->tile), TY = hp::stencil2d_tile_h(b->tile);
    for (auto& kv : b->by_stencil) {
        StencilGroup gr{kv.first.first, kv.first.second, {}};
        std::vector<int4> it;
        for (uint32_t i : kv.second) {
            const hp::Geo& g = b->members[i].plan.g;
            const int nt = ((g.W + TX - 1) / TX) * ((g.H + TY - 1) / TY);
            for (int t = 0; t < nt; t++) it.push_back(make_int4((int)i, t, 0, 0));
        }
        STRY(bupload(b, &gr.items.d, it, L));
        gr.items.n = (int)it.size();
        b->stencil_groups.push_back(gr);
    }
    // the persistent step's phases, when every member's sides are among its lengths
    b->persistent_ok = true;
    for (const HostMember& hm : b->members)
        if (!hp::batch_step_persistent_size(hm.plan.g.W) || !hp::batch_step_persistent_size(hm.plan.g.H) ||
            hm.plan.g.D != 1)
            b->persistent_ok = false;
    b->pitems = ItemList{};
    const int nseg = hp::batch_step_persistent_segments();
    if (b->persistent_ok && nseg + 1 <= (int)(sizeof(b->pstart) / sizeof(b->pstart[0]))) {
        std::vector<std::vector<int4>> seg(nseg);
        const int PTX = hp::batch_step_persistent_tile_w(), PTY = hp::batch_step_persistent_tile_h();
        for (uint32_t i = 0; i < (uint32_t)b->members.size(); i++) {
            const hp::MemberPlan& p = b->members[i].plan;
            const HostMember& hm = b->members[i];
            const int pairs = (p.g.H + 1) / 2;
            const int CB = hp::batch_step_persistent_block_lines(p.g.H);
            if (p.K > 0) {
                auto& rf = seg[hp::batch_step_persistent_segment(0, p.g.W, p.C, false)];
                for (int sl = 0; sl < (int)p.src_chan.size(); sl++)
                    for (int pr = 0; pr < pairs; pr++) rf.push_back(make_int4((int)i, pr, sl, 0));
                auto& sp = seg[hp::batch_step_persistent_segment(1, p.g.H, p.C, false)];
                for (int wi = 0; wi < (int)p.spec_work.size(); wi++)
                    for (int blk = 0; blk < (p.g.Wh + CB - 1) / CB; blk++) sp.push_back(make_int4((int)i, blk, wi, 0));
                auto& ri = seg[hp::batch_step_persistent_segment(2, p.g.W, p.C, false)];
                for (int wi = 0; wi < (int)p.tgt_work.size(); wi++)
                    for (int pr = 0; pr < pairs; pr++) ri.push_back(make_int4((int)i, pr, wi, 0));
            }
            const float s = hm.cfg.flow.distribution_size, ta = (2.0f * s) * (2.0f * s);
            int e;
            const bool pow2 = frexpf(ta, &e) == 0.5f && ta >= 1.17549435e-38f && ta < INFINITY;
            auto& st = seg[hp::batch_step_persistent_segment(3, p.g.W, p.C, pow2)];
            const int nt = ((p.g.W + PTX - 1) / PTX) * ((p.g.H + PTY - 1) / PTY);
            for (int t = 0; t < nt; t++) st.push_back(make_int4((int)i, t, 0, 0));
        }
        std::vector<int4> all;
        for (int k = 0; k < nseg; k++) {
            b->pstart[k] = (int)all.size();
            all.insert(all.end(), seg[k].begin(), seg[k].end());
        }
        b->pstart[nseg] = (int)all.size();
        STRY(bupload(b, &b->pitems.d, all, L));
        b->pitems.n = (int)all.size();
    } else {
        b->persistent_ok = false;
    }
    if (!b->persistent_ok) b->persistent = false;
    return HP_OK;
}

// Switches between the multi-launch step and the persistent step; the graphs are recaptured.
hp_status batch_set_step_mode(hp_batch* b, int mode) {
    if (mode != HP_STEP_MULTI_LAUNCH && mode != HP_STEP_PERSISTENT) return b->fail(HP_ERR_SHAPE, "unknown step mode");
    if (mode == HP_STEP_PERSISTENT) {
        if (!b->persistent_ok)
            return b->fail(HP_ERR_UNSUPPORTED_SIZE,
                           "the persistent step covers member sides 64, 128, 192 and 256 only");
        if (b->fv.rp != 0) return b->fail(HP_ERR_SHAPE, "the persistent step runs radix order 0");
    }
    BTRY(b, hipStreamSynchronize(b->stream));
    for (int p = 0; p < 2; p++)
        if (b->graph[p]) {
            (void)hipGraphExecDestroy(b->graph[p]);
            b->graph[p] = nullptr;
        }
    b->persistent = mode == HP_STEP_PERSISTENT;
    return build_graphs(b);
}

void release(hp_batch* b) {
    if (!b) return;
    for (int p = 0; p < 2; p++)
        if (b->graph[p]) (void)hipGraphExecDestroy(b->graph[p]);
    for (void* p : b->tape_allocs) (void)hipFree(p);
    for (void* p : b->item_allocs) (void)hipFree(p);
    for (void* p : b->allocs) (void)hipFree(p);
    if (b->h_stage) (void)hipHostFree(b->h_stage);
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
        if (st == HP_OK && b->members[i].plan.direct) {
            // the batched passes and the gradient tape run on the FFT; such members step on
            // their own contexts
            st = HP_ERR_UNSUPPORTED_SIZE;
            perr = "population members need sides 2^a * 3^b between 8 and 4096";
        }
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
        for (const hp_kernel& kc : hm.kernels) hm.radius.push_back((int)hp::actual_radius(kc.radius, cfg->kernel_radius));
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
    std::map<std::pair<int, int>, std::vector<uint32_t>>& by_size = b->by_size;
    std::map<std::pair<int, bool>, std::vector<uint32_t>>& by_stencil = b->by_stencil;
    std::vector<int4> copy_it, tiles;
    const int ATX = hp::adj_tile_w(), ATY = hp::adj_tile_h();
    for (uint32_t i = 0; i < count; i++) {
        HostMember& hm = b->members[i];
        const hp::MemberPlan& p = hm.plan;
        hp::MemberDesc& d = b->desc[i];
        memset(&d, 0, sizeof(d));
        const size_t CN = (size_t)p.C * p.g.N;
        TRY(balloc(b, &d.A[0], CN));
        TRY(balloc(b, &d.A[1], CN));
        TRY(balloc(b, &d.U, CN));
        TRY(balloc(b, &d.snap, CN));
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
        const int nparts = (int)std::max<size_t>(1, std::min<size_t>(256, CN / 4096));
        for (int q = 0; q < nparts; q++) copy_it.push_back(make_int4((int)i, q, nparts, 0));
        const int nt = ((p.g.W + ATX - 1) / ATX) * ((p.g.H + ATY - 1) / ATY);
        for (int t = 0; t < nt; t++) tiles.push_back(make_int4((int)i, t, 0, 0));
    }
    TRY(bupload(b, &b->d_desc, b->desc));
    TRY(bupload(b, &b->copy_items.d, copy_it));
    b->copy_items.n = (int)copy_it.size();
    TRY(bupload(b, &b->adj_tiles.d, tiles));
    b->adj_tiles.n = (int)tiles.size();
    TRY(build_items(b));
    size_t fused = 0;
    for (auto& kv : by_stencil) fused += kv.second.size();
    if (fused != count) {
        g_batch_error = "every population member must suit the fused 2D stencil (sides >= 5, 0 < s <= 1)";
        release(b);
        return HP_ERR_UNSUPPORTED_SIZE;
    }
    TRY(balloc(b, &b->d_flags, 4));
    TRY(balloc(b, &b->d_bar, 2));
    BTRY(b, hipMemset(b->d_bar, 0, 2 * sizeof(unsigned)));
    TRY(balloc(b, &b->box_lo, max_cn));
    TRY(balloc(b, &b->box_hi, max_cn));
    TRY(balloc(b, &b->d_moments, 3 * (size_t)count));
    TRY(balloc(b, &b->d_partials, (size_t)kMassParts * max_c));
    TRY(balloc(b, &b->d_mass, (size_t)max_c + 1));
    BTRY(b, hipMemset(b->d_flags, 0, 16));
    BTRY(b, hipDeviceSynchronize());
    TRY(build_graphs(b));
#undef TRY
    *out = b;
    return HP_OK;
}

hp_status batch_set_plan(hp_batch* b, const hp_plan& p) {
    if (p.radix_order < 0 || p.radix_order > 1 || p.column_split < 1 || p.column_split > 2 || p.stencil_tile < 0 ||
        p.stencil_tile >= hp::stencil2d_tile_count())
        return b->fail(HP_ERR_SHAPE, "plan field out of range");
    if (b->persistent && p.radix_order != 0) return b->fail(HP_ERR_SHAPE, "the persistent step runs radix order 0");
    for (const HostMember& hm : b->members)
        if (!hp::stencil2d_supported(hm.plan.g, hm.cfg.flow.distribution_size, p.stencil_tile))
            return b->fail(HP_ERR_UNSUPPORTED_SIZE, "stencil tile does not fit every member's grid");
    BTRY(b, hipStreamSynchronize(b->stream));
    b->fv.rp = p.radix_order;
    b->fv.cs = p.column_split;
    b->tile = p.stencil_tile;
    STRY(build_items(b));
    return build_graphs(b);  // captured now, not in the first advance after the change
}

hp_plan batch_plan(const hp_batch* b) {
    hp_plan p{};
    p.fft_path = 1;
    p.spectral_group = 1;
    p.radix_order = b->fv.rp;
    p.column_split = b->fv.cs;
    p.stencil_tile = b->tile;
    return p;
}

// Coordinate search as for a single context: the four LDS FFT variants with the current
// tile, then the tiles with the fastest variant.  Every member's state is restored after.
hp_status batch_autotune(hp_batch* b, uint32_t steps, hp_plan* best_out, uint32_t maxr, hp_plan* tried,
                         double* us_out, uint32_t* ntried) {
    if (steps == 0) steps = 20;
    const size_t M = b->members.size();
    std::vector<float> t0(M);
    std::vector<uint64_t> s0(M);
    for (size_t i = 0; i < M; i++) {
        t0[i] = b->members[i].time;
        s0[i] = b->members[i].step;
    }
    const uint64_t g0 = b->general_steps;
    const uint32_t r0 = b->max_radius;
    const bool had_tape = b->tape_recorded;
    // every member's state into one temporary buffer (advance() uses the snapshot slots itself)
    size_t total = 0;
    std::vector<size_t> off(M);
    for (size_t i = 0; i < M; i++) {
        off[i] = total;
        total += (size_t)b->desc[i].C * b->desc[i].g.N;
    }
    float* keep = nullptr;
    BTRY(b, hipMalloc((void**)&keep, std::max<size_t>(1, total) * sizeof(float)));
    auto save = [&](bool back) -> hp_status {
        for (size_t i = 0; i < M; i++) {
            const hp::MemberDesc& m = b->desc[i];
            const size_t bytes = (size_t)m.C * m.g.N * sizeof(float);
            if (back) BTRY(b, hipMemcpyAsync(m.A[b->cur], keep + off[i], bytes, hipMemcpyDeviceToDevice, b->stream));
            else BTRY(b, hipMemcpyAsync(keep + off[i], m.A[b->cur], bytes, hipMemcpyDeviceToDevice, b->stream));
        }
        return HP_OK;
    };
    hp_status sst = save(false);
    if (sst != HP_OK) {
        (void)hipFree(keep);
        return sst;
    }
    hipEvent_t ea = nullptr, eb = nullptr;
    BTRY(b, hipEventCreate(&ea));
    BTRY(b, hipEventCreate(&eb));
    std::vector<double> ref, mom(3 * M);
    uint32_t n = 0;
    double best_us = INFINITY;
    const hp_plan start = batch_plan(b);
    hp_plan best = start;
    auto measure = [&](const hp_plan& p) -> hp_status {
        if (b->persistent && p.radix_order != 0) return HP_OK;  // the persistent step runs radix order 0
        const hp_status ps = batch_set_plan(b, p);
        if (ps == HP_ERR_UNSUPPORTED_SIZE) return HP_OK;  // a candidate these members cannot take
        if (ps != HP_OK) return ps;  // anything else leaves the plan unusable: stop tuning
        STRY(save(true));
        STRY(advance(b, 2));
        STRY(save(true));
        BTRY(b, hipEventRecord(ea, b->stream));
        STRY(advance(b, steps));
        BTRY(b, hipEventRecord(eb, b->stream));
        BTRY(b, hipEventSynchronize(eb));
        float ms = 0.0f;
        BTRY(b, hipEventElapsedTime(&ms, ea, eb));
        hp::launch_batch_moments(b->stream, b->d_desc, (int)M, b->cur, b->d_moments);
        BTRY(b, hipMemcpyAsync(mom.data(), b->d_moments, mom.size() * sizeof(double), hipMemcpyDeviceToHost,
                               b->stream));
        BTRY(b, hipStreamSynchronize(b->stream));
        bool ok = true;
        if (ref.empty()) {
            ref = mom;
        } else {
            for (size_t j = 0; j < mom.size() && ok; j++)
                ok = fabs(mom[j] - ref[j]) <= 1e-4 * std::max(1.0, fabs(ref[j]));
        }
        const double us = 1e3 * ms / steps;
        if (n < maxr) {
            if (tried) tried[n] = p;
            if (us_out) us_out[n] = ok ? us : -us;
        }
        n++;
        if (ok && us < best_us) {
            best_us = us;
            best = p;
        }
        return HP_OK;
    };
    hp_status st = measure(start);
    for (int rp = 0; rp <= 1 && st == HP_OK; rp++)
        for (int cs = 1; cs <= 2 && st == HP_OK; cs++) {
            hp_plan p = start;
            p.radix_order = rp;
            p.column_split = cs;
            if (rp == start.radix_order && cs == start.column_split) continue;
            st = measure(p);
        }
    const hp_plan fft_best = best;
    for (int t = 0; t < hp::stencil2d_tile_count() && st == HP_OK; t++) {
        if (t == fft_best.stencil_tile) continue;
        hp_plan p = fft_best;
        p.stencil_tile = t;
        st = measure(p);
    }
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    if (st == HP_OK) st = batch_set_plan(b, best);
    if (st == HP_OK) st = save(true);
    if (st == HP_OK && hipStreamSynchronize(b->stream) != hipSuccess) st = b->fail(HP_ERR_HIP, "synchronise");
    (void)hipFree(keep);
    if (st != HP_OK) return st;
    for (size_t i = 0; i < M; i++) {
        b->members[i].time = t0[i];
        b->members[i].step = s0[i];
    }
    b->general_steps = g0;
    b->max_radius = r0;
    (void)had_tape;
    b->tape_recorded = false;  // tuning stepped the population: a tape must be recorded again
    if (best_out) *best_out = best;
    if (ntried) *ntried = n;
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

// ---- fitness evaluation (FitnessEvaluator::evaluate, fitness.rs:26-86) ----------------------
// Device buffers of one evaluation, freed when it ends.
struct EvalBuffers {
    std::vector<void*> p;
    ~EvalBuffers() {
        for (void* q : p) (void)hipFree(q);
    }
    template <typename T>
    hp_status get(hp_batch* b, T** out, size_t n) {
        void* v = nullptr;
        BTRY(b, hipMalloc(&v, std::max<size_t>(1, n) * sizeof(T)));
        p.push_back(v);
        *out = (T*)v;
        return HP_OK;
    }
};

// One sample of a member in the reference's f32 terms (record_sample, fitness.rs:146-160;
// centre of mass and radius, fitness.rs:196-244), from its FP64 sums.
struct FitSample {
    float cx = 0, cy = 0, radius = 0, mass = 0, maxv = 0;
    uint32_t active = 0;
    double sumsq = 0;
};

FitSample fit_sample(const double* st, int W, int H) {
    FitSample o;
    o.mass = (float)st[0];
    const double pos = st[1];
    if ((float)pos > 1e-6f) {
        o.cx = (float)(st[2] / pos);
        o.cy = (float)(st[3] / pos);
        // second moment about the f32 centre: sum m ((x - cx)^2 + (y - cy)^2) expanded
        const double cx = o.cx, cy = o.cy;
        const double mom = st[4] - 2.0 * cx * st[2] - 2.0 * cy * st[3] + (cx * cx + cy * cy) * pos;
        o.radius = (float)sqrt(std::max(0.0, mom) / pos);
    } else {
        o.cx = (float)W / 2.0f;
        o.cy = (float)H / 2.0f;
        o.radius = 0.0f;
    }
    o.maxv = (float)st[5];
    o.active = (uint32_t)st[6];
    o.sumsq = st[8];
    return o;
}

float clamp01f(float v) { return v != v ? v : std::min(std::max(v, 0.0f), 1.0f); }

// compute_state_similarity (fitness.rs:353-366) from a dot product and two sums of squares.
float similarity_of(double dot, double qa, double qb) {
    const float na = (float)sqrt(qa), nb = (float)sqrt(qb);
    if (na > 1e-6f && nb > 1e-6f) return (float)(dot / (sqrt(qa) * sqrt(qb)));
    return 0.0f;
}

struct MemberEval {
    int W = 0, H = 0;
    float initial_mass = 0;
    std::vector<FitSample> s;    // samples 0 .. S-1
    std::vector<uint64_t> step;  // their steps
    const double* fin = nullptr; // final-state sums
    double init_sumsq = 0;
    std::map<std::pair<int, int>, double> dot;  // (slot a, slot b) -> dot product; slot 0 = initial
};

float displacement(const MemberEval& e) {
    if (e.s.size() < 2) return 0.0f;
    const float dx = e.s.back().cx - e.s[0].cx, dy = e.s.back().cy - e.s[0].cy;
    return sqrtf(dx * dx + dy * dy);
}

float m_persistence(const MemberEval& e) {  // fitness.rs:271-291
    if (e.initial_mass < 1e-6f) return 0.0f;
    const float final_mass = e.s.empty() ? 0.0f : e.s.back().mass;
    const float ratio = final_mass / e.initial_mass;
    const float fmax_ = e.s.empty() ? 0.0f : e.s.back().maxv;
    float conc = 0.0f;
    if (final_mass > 1e-6f) conc = fmax_ / (final_mass / (float)(e.s.empty() ? 1u : e.s.back().active));
    return clamp01f(ratio * std::min(conc, 10.0f) / 10.0f);
}

float m_compactness(const MemberEval& e) {  // fitness.rs:294-306
    const float r = e.s.empty() ? 3.40282347e38f : e.s.back().radius;
    const float maxr = (float)std::min(e.W, e.H) / 2.0f;
    return clamp01f(1.0f - r / maxr);
}

float m_locomotion(const MemberEval& e) {  // fitness.rs:309-326
    if (e.s.size() < 2) return 0.0f;
    const float maxd = sqrtf((float)((uint64_t)e.W * e.W + (uint64_t)e.H * e.H));
    return clamp01f(displacement(e) / maxd);
}

float m_periodicity(const MemberEval& e, uint64_t period, float tol) {  // fitness.rs:329-350
    for (size_t j = 0; j < e.s.size(); j++)
        if (e.step[j] == period) {
            const float sim = similarity_of(e.dot.at({0, (int)j + 1}), e.init_sumsq, e.s[j].sumsq);
            if (sim > 1.0f - tol) return sim;
        }
    float best = 0.0f;
    for (size_t j = 0; j < e.s.size(); j++)
        if (e.step[j] > 0) best = std::max(best, similarity_of(e.dot.at({0, (int)j + 1}), e.init_sumsq, e.s[j].sumsq));
    return best;
}

float m_complexity(const MemberEval& e) {  // fitness.rs:369-398
    const double n = (double)e.W * e.H;
    const float mean = (float)(e.fin[0] / n);
    if (mean < 1e-6f) return 0.0f;
    const double md = e.fin[0] / n;
    const float var = (float)std::max(0.0, e.fin[8] / n - md * md);
    const float gmean = (float)(e.fin[9] / (2.0 * n));
    return std::min((sqrtf(var) + gmean) / 2.0f, 1.0f);
}

float m_mass_concentration(const MemberEval& e) {  // fitness.rs:401-413
    const float mean = (float)(e.fin[0] / ((double)e.W * e.H));
    if (mean < 1e-6f) return 0.0f;
    return std::min(((float)e.fin[5] / mean) / 100.0f, 1.0f);
}

float m_glider(const MemberEval& e, float min_disp) {  // fitness.rs:416-453
    const float loco = m_locomotion(e), comp = m_compactness(e);
    if (displacement(e) < min_disp) return 0.0f;
    float cons = 0.5f;
    if (e.s.size() >= 2) {  // f32 sums in sample order, as the reference's iterators
        float sum = 0.0f;
        for (const FitSample& x : e.s) sum += x.radius;
        const float mean = sum / (float)e.s.size();
        float v = 0.0f;
        for (const FitSample& x : e.s) {
            const float d = x.radius - mean;
            v += d * d;
        }
        v /= (float)e.s.size();
        const float cv = sqrtf(v) / std::max(mean, 1e-6f);
        cons = clamp01f(1.0f - cv);
    }
    return loco * comp * cons * m_persistence(e);
}

float m_oscillator(const MemberEval& e, uint64_t max_period, float thr) {  // fitness.rs:456-483
    float best = 0.0f;
    for (uint64_t period = 1; period <= max_period; period++)
        for (size_t i = 0; i < e.s.size(); i++)
            for (size_t j = i + 1; j < e.s.size(); j++)
                if (e.step[j] - e.step[i] == period) {
                    const float sim = similarity_of(e.dot.at({(int)i + 1, (int)j + 1}), e.s[i].sumsq, e.s[j].sumsq);
                    if (sim > thr) best = std::max(best, sim * (1.0f - (float)period / (float)max_period * 0.5f));
                }
    return best;
}

float m_stability(const MemberEval& e) {  // fitness.rs:486-511
    const double n = (double)e.W * e.H;
    const float neg = std::min((float)e.fin[7] / (float)n, 1.0f);
    const float mx = (float)e.fin[5];
    const float extreme = mx > 10.0f ? (mx - 10.0f) / mx : 0.0f;
    const float final_mass = e.s.empty() ? 0.0f : e.s.back().mass;
    const float err = e.initial_mass > 1e-6f ? fabsf((final_mass - e.initial_mass) / e.initial_mass) : 0.0f;
    return clamp01f(1.0f - neg - extreme - err);
}

float metric_score(const hp_metric& mt, const MemberEval& e) {
    switch (mt.kind) {
    case HP_METRIC_PERSISTENCE: return m_persistence(e);
    case HP_METRIC_COMPACTNESS: return m_compactness(e);
    case HP_METRIC_LOCOMOTION: return m_locomotion(e);
    case HP_METRIC_PERIODICITY: return m_periodicity(e, mt.period, mt.tolerance);
    case HP_METRIC_COMPLEXITY: return m_complexity(e);
    case HP_METRIC_MASS_CONCENTRATION: return m_mass_concentration(e);
    case HP_METRIC_GLIDER_SCORE: return m_glider(e, mt.min_displacement);
    case HP_METRIC_OSCILLATOR_SCORE: return m_oscillator(e, mt.max_period, mt.threshold);
    case HP_METRIC_STABILITY: return m_stability(e);
    default: return 0.0f;  // Custom: scored outside the evaluator
    }
}

hp_status evaluate(hp_batch* b, const hp_eval_config& ev, const hp_metric* metrics, uint32_t nm, bool normalize,
                   float* fitness, float* scores, hp_behavior* behavior, hp_sample* samples) {
    for (uint32_t k = 0; k < nm; k++)
        if (metrics[k].kind < 0 || metrics[k].kind > HP_METRIC_CUSTOM)
            return b->fail(HP_ERR_SHAPE, "unknown metric kind " + std::to_string(metrics[k].kind));
    const size_t M = b->members.size();
    const uint64_t every = std::max<uint64_t>(1, ev.sample_interval);
    const uint64_t S = 1 + ev.steps / every;
    if (S > 100000) return b->fail(HP_ERR_SHAPE, "more than 100000 samples");
    bool need_snap = false;
    uint64_t max_period = 0;
    for (uint32_t k = 0; k < nm; k++) {
        if (metrics[k].kind == HP_METRIC_PERIODICITY) need_snap = true;
        if (metrics[k].kind == HP_METRIC_OSCILLATOR_SCORE) {
            need_snap = true;
            max_period = std::max(max_period, metrics[k].max_period);
        }
    }
    EvalBuffers eb;
    // work items: per member parts of at least 8192 cells, at most 32
    std::vector<int4> items;
    std::vector<int> first(M + 1, 0);
    for (size_t i = 0; i < M; i++) {
        first[i] = (int)items.size();
        const long long n = (long long)b->desc[i].g.W * b->desc[i].g.H;
        const int parts = (int)std::max<long long>(1, std::min<long long>(32, n / 8192));
        for (int q = 0; q < parts; q++) items.push_back(make_int4((int)i, q, parts, 0));
    }
    first[M] = (int)items.size();
    int4* d_items = nullptr;
    int* d_first = nullptr;
    double *d_part = nullptr, *d_stats = nullptr;
    const size_t rows = (size_t)S + 2;  // initial, S samples, final
    STRY(eb.get(b, &d_items, items.size()));
    STRY(eb.get(b, &d_first, first.size()));
    STRY(eb.get(b, &d_part, items.size() * hp::kFitStats));
    STRY(eb.get(b, &d_stats, rows * M * hp::kFitStats));
    BTRY(b, hipMemcpy(d_items, items.data(), items.size() * sizeof(int4), hipMemcpyHostToDevice));
    BTRY(b, hipMemcpy(d_first, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice));
    // snapshot slots: 0 the initial channel sum, 1 .. S the samples
    float** d_snap_base = nullptr;
    if (need_snap) {
        std::vector<float*> base(M);
        for (size_t i = 0; i < M; i++) {
            const size_t n = (size_t)b->desc[i].g.W * b->desc[i].g.H;
            STRY(eb.get(b, &base[i], (S + 1) * n));
        }
        STRY(eb.get(b, &d_snap_base, M));
        BTRY(b, hipMemcpy(d_snap_base, base.data(), M * sizeof(float*), hipMemcpyHostToDevice));
    }
    auto sample = [&](size_t row, int slot, bool grad) -> hp_status {
        hp::launch_fit_stats(b->stream, b->d_desc, d_items, (int)items.size(), b->cur, grad,
                             slot >= 0 ? d_snap_base : nullptr, slot, d_part);
        hp::launch_fit_reduce(b->stream, d_first, (int)M, d_part, d_stats + row * M * hp::kFitStats);
        BTRY(b, hipGetLastError());
        return HP_OK;
    };
    b->tape_recorded = false;
    STRY(sample(0, need_snap ? 0 : -1, false));  // EvaluationTrajectory::new, before the warmup
    STRY(advance(b, ev.warmup_steps));
    STRY(sample(1, need_snap ? 1 : -1, false));
    uint64_t done = 0;
    for (uint64_t j = 1; j < S; j++) {
        STRY(advance(b, every));
        done += every;
        STRY(sample(1 + j, need_snap ? (int)(1 + j) : -1, false));
    }
    STRY(advance(b, ev.steps - done));
    STRY(sample(S + 1, -1, true));
    // similarity dot products: periodicity (initial vs each sample), oscillator (sample pairs
    // whose steps differ by 1 .. max_period)
    std::vector<int4> pairs;
    if (need_snap)
        for (size_t i = 0; i < M; i++) {
            for (uint64_t j = 0; j < S; j++) pairs.push_back(make_int4((int)i, 0, (int)(1 + j), 0));
            for (uint64_t a = 0; a < S; a++)
                for (uint64_t c = a + 1; c < S && (c - a) * every <= max_period; c++)
                    pairs.push_back(make_int4((int)i, (int)(1 + a), (int)(1 + c), 0));
        }
    std::vector<double> dots(pairs.size());
    if (!pairs.empty()) {
        int4* d_pairs = nullptr;
        double* d_dots = nullptr;
        STRY(eb.get(b, &d_pairs, pairs.size()));
        STRY(eb.get(b, &d_dots, pairs.size()));
        BTRY(b, hipMemcpyAsync(d_pairs, pairs.data(), pairs.size() * sizeof(int4), hipMemcpyHostToDevice, b->stream));
        hp::launch_fit_dots(b->stream, b->d_desc, d_pairs, (int)pairs.size(), d_snap_base, d_dots);
        BTRY(b, hipGetLastError());
        BTRY(b, hipMemcpyAsync(dots.data(), d_dots, dots.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    }
    std::vector<double> st(rows * M * hp::kFitStats);
    BTRY(b, hipMemcpyAsync(st.data(), d_stats, st.size() * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    BTRY(b, hipStreamSynchronize(b->stream));
    std::vector<MemberEval> E(M);
    for (size_t i = 0; i < M; i++) {
        MemberEval& e = E[i];
        e.W = b->desc[i].g.W;
        e.H = b->desc[i].g.H;
        const double* r0 = &st[(0 * M + i) * hp::kFitStats];
        e.initial_mass = (float)r0[0];
        e.init_sumsq = r0[8];
        for (uint64_t j = 0; j < S; j++) {
            e.s.push_back(fit_sample(&st[((1 + j) * M + i) * hp::kFitStats], e.W, e.H));
            e.step.push_back(j * every);
        }
        e.fin = &st[((S + 1) * M + i) * hp::kFitStats];
    }
    for (size_t q = 0; q < pairs.size(); q++) E[pairs[q].x].dot[{pairs[q].y, pairs[q].z}] = dots[q];
    float total_w = 0.0f;
    for (uint32_t k = 0; k < nm; k++) total_w += metrics[k].weight;
    for (size_t i = 0; i < M; i++) {
        const MemberEval& e = E[i];
        float comb = 0.0f;
        for (uint32_t k = 0; k < nm; k++) {
            const float sc = metric_score(metrics[k], e);
            if (scores) scores[i * nm + k] = sc;
            comb += (normalize ? clamp01f(sc) : sc) * metrics[k].weight;
        }
        if (fitness) fitness[i] = comb / std::max(total_w, 1e-6f);
        if (behavior) {  // to_behavior_stats (fitness.rs:163-191)
            hp_behavior& bh = behavior[i];
            bh.final_mass = e.s.back().mass;
            bh.initial_mass = e.initial_mass;
            bh.total_displacement = displacement(e);
            bh.final_radius = e.s.back().radius;
            bh.active_cells = e.s.back().active;
            float mx = 0.0f;
            for (const FitSample& x : e.s) mx = std::max(mx, x.maxv);
            bh.max_activation = mx;
        }
        if (samples)
            for (uint64_t j = 0; j < S; j++) {
                hp_sample& o = samples[i * S + j];
                const FitSample& x = e.s[j];
                o = hp_sample{x.cx, x.cy, x.radius, x.mass, x.maxv, x.active};
            }
    }
    return HP_OK;
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
        // every channel is checked before any is copied: a refused state leaves the member as it was
        for (int c = 0; c < m.C; c++) {
            if (!channels[c]) return b->fail(HP_ERR_NULL, "channel pointer is NULL");
            for (long long j = 0; j < m.g.N; j++)
                if (!isfinite(channels[c][j])) return b->fail(HP_ERR_NONFINITE, "input state holds NaN/Inf");
        }
        for (int c = 0; c < m.C; c++)
            BTRY(b, hipMemcpyAsync(m.A[b->cur] + (size_t)c * m.g.N, channels[c], m.g.N * sizeof(float),
                                   hipMemcpyHostToDevice, b->stream));
        BTRY(b, hipStreamSynchronize(b->stream));
        b->members[member].time = 0.0f;
        b->members[member].step = 0;
        b->tape_recorded = false;
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
    return guarded(b, [&]() -> hp_status {
        b->tape_recorded = false;
        return advance(b, steps);
    });
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

hp_status hp_batch_set_step_mode(hp_batch* b, int32_t mode) {
    return guarded(b, [&]() -> hp_status { return batch_set_step_mode(b, mode); });
}

hp_status hp_batch_step_mode(const hp_batch* b, int32_t* mode, uint64_t* barrier_timeouts) {
    if (!b) return HP_ERR_NULL;
    if (mode) *mode = b->persistent ? HP_STEP_PERSISTENT : HP_STEP_MULTI_LAUNCH;
    if (barrier_timeouts) *barrier_timeouts = b->barrier_timeouts;
    return HP_OK;
}

hp_status hp_batch_theta_count(const hp_batch* b, uint32_t member, uint32_t* count) {
    if (!b || !count) return HP_ERR_NULL;
    if (member >= b->members.size()) return HP_ERR_SHAPE;
    *count = (uint32_t)theta_count(b->members[member]);
    return HP_OK;
}

hp_status hp_batch_get_plan(const hp_batch* b, hp_plan* out) {
    if (!b || !out) return HP_ERR_NULL;
    *out = batch_plan(b);
    return HP_OK;
}

hp_status hp_batch_set_plan(hp_batch* b, const hp_plan* plan) {
    if (!plan) return HP_ERR_NULL;
    return guarded(b, [&]() -> hp_status { return batch_set_plan(b, *plan); });
}

hp_status hp_batch_autotune(hp_batch* b, uint32_t steps, hp_plan* best, uint32_t max_results, hp_plan* tried,
                            double* us_per_step, uint32_t* n_tried) {
    return guarded(b, [&]() -> hp_status {
        return batch_autotune(b, steps, best, max_results, tried, us_per_step, n_tried);
    });
}

}  // extern "C"

hp_status hp::batch_device_state(hp_batch* b, uint32_t member, hp::DeviceState* out) {
    if (!b || !out) return HP_ERR_NULL;
    if (member >= b->members.size()) return b->fail(HP_ERR_SHAPE, "member index out of range");
    const hp::MemberDesc& m = b->desc[member];
    out->data = m.A[b->cur];
    out->floats = (size_t)m.C * (size_t)m.g.N;
    out->stream = b->stream;
    out->device = b->device;
    out->W = (uint32_t)m.g.W;
    out->H = (uint32_t)m.g.H;
    out->D = 1;
    out->C = (uint32_t)m.C;
    return HP_OK;
}

extern "C" {

hp_status hp_eval_sample_count(const hp_eval_config* ev, uint32_t* count) {
    if (!ev || !count) return HP_ERR_NULL;
    *count = (uint32_t)(1 + ev->steps / std::max<uint64_t>(1, ev->sample_interval));
    return HP_OK;
}

hp_status hp_batch_evaluate(hp_batch* b, const hp_eval_config* ev, const hp_metric* metrics, uint32_t num_metrics,
                            int normalize, float* fitness, float* scores, hp_behavior* behavior,
                            hp_sample* samples) {
    return guarded(b, [&]() -> hp_status {
        if (!ev || (num_metrics && !metrics)) return b->fail(HP_ERR_NULL, "evaluation or metric list is NULL");
        return evaluate(b, *ev, metrics, num_metrics, normalize != 0, fitness, scores, behavior, samples);
    });
}

hp_status hp_batch_tape_forward(hp_batch* b, uint64_t steps, uint32_t checkpoint_every, uint32_t* used_every) {
    return guarded(b, [&]() -> hp_status {
        hp_status st = tape_forward(b, steps, checkpoint_every);
        if (st == HP_OK && used_every) *used_every = b->tape_k;
        return st;
    });
}

hp_status hp_batch_tape_backward(hp_batch* b, const float* const* dl_dfinal, float* const* dl_dinitial,
                                 double* dtheta) {
    return guarded(b, [&]() -> hp_status {
        if (!dl_dfinal) return b->fail(HP_ERR_NULL, "dl_dfinal is NULL");
        return tape_backward(b, dl_dfinal, dl_dinitial, dtheta);
    });
}

}  // extern "C"
