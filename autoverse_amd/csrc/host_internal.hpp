// host_internal.hpp -- host-side helpers shared by the library's C++ sources.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/flowlenia_hip.h"
#include "hp_common.hpp"

namespace hp {

// Everything a propagator needs on the device, computed on the host from a config:
// geometry, folded kernel spectra, source slots and the work lists of the passes.
struct MemberPlan {
    Geo g{};
    int C = 0, K = 0;
    bool is3d = false;
    bool wave = false;  // 2D sides 256..2048: column-major spectra, wave-level FFTs
    bool grouped = true;  // spectral work: one item per source (else one per kernel)
    std::vector<float> kq;
    std::vector<int> src_chan;  // source slot -> channel
    std::vector<SpecWork> spec_work;
    std::vector<int> klist_spec;
    std::vector<TgtWork> tgt_work;
    std::vector<int> klist_tgt;
    std::vector<int> tgt_chan;
    std::vector<GrowthParams> gp;
    // sides the FFT passes do not cover: the convolutions are summed directly
    bool direct = false;
    std::vector<DirectKernel> dk;
    std::vector<float> taps;
    std::vector<int2> cols;
    int rmax = 0;
};

// group: 1 = one spectral item per source, 0 = one per kernel, -1 = choose.
// allow_wave: use the column-major wave path when the sizes allow it.
hp_status plan_member(const hp_config* cfg, bool allow_wave, int group, MemberPlan* out, std::string* err);

hp_status validate_config(const hp_config* cfg, uint32_t* bad_kernel, uint32_t* bad_channel);
// The message hp_last_global_error returns on this thread (calls that have no context).
void set_global_error(const std::string& msg);

// The device-resident state of a context or of one population member: C*N floats laid out
// [C][D*H*W] (the order of an FLWA frame), the stream its steps run on, and its shape.
struct DeviceState {
    const float* data = nullptr;
    size_t floats = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    uint32_t W = 0, H = 0, D = 0, C = 0;
};
hp_status ctx_device_state(hp_ctx* ctx, DeviceState* out);

// A context's device buffers and step bookkeeping, for a caller that drives its own step on
// them (the embedded-parameter context): the state ping-pong A[*cur], the general-path boxes,
// the snapshot buffer, and the time and step the context reports.
struct CtxView {
    hipStream_t stream = nullptr;
    int device = 0;
    Geo g{};
    int C = 0, K = 0;
    float* A[2] = {nullptr, nullptr};
    int* cur = nullptr;
    float* time = nullptr;
    uint64_t* step = nullptr;
    float4* box_lo = nullptr;
    float4* box_hi = nullptr;
    float* snapshot = nullptr;
};
// A context whose convolutions run on the LDS FFT passes (the path ctx_enqueue_conv needs).
hp_status ctx_create_lds(const hp_config* cfg, int device, hp_ctx** out);
hp_status ctx_view(hp_ctx* ctx, CtxView* out);
// Every kernel's convolution of its source channel of A (2D) into conv + k*N (no growth).
hp_status ctx_enqueue_conv(hp_ctx* ctx, const float* A, float* conv);
// Message of the context's last error (also what hp_last_error returns).
const char* ctx_error(const hp_ctx* ctx);
hp_status batch_device_state(hp_batch* b, uint32_t member, DeviceState* out);
uint32_t actual_radius(float radius, uint32_t kernel_radius);
std::vector<float> build_ring_kernel(const hp_kernel& k, const hp_ring* rings, uint32_t kernel_radius, int dim,
                                     uint32_t* size_out);
std::vector<float> folded_spectrum(const std::vector<float>& k, uint32_t size, int dim, uint32_t W, uint32_t H,
                                   uint32_t D);
hp_status seed_generate(const hp_seed* seed, uint32_t w, uint32_t h, uint32_t d, uint32_t c, float* const* out);

}  // namespace hp
