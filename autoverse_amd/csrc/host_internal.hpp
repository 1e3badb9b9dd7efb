// host_internal.hpp -- host-side helpers shared by the library's C++ sources.
#pragma once

#include <stdint.h>

#include <vector>

#include "../../include/flowlenia_hip.h"

namespace hp {

hp_status validate_config(const hp_config* cfg, uint32_t* bad_kernel, uint32_t* bad_channel);
uint32_t actual_radius(float radius, uint32_t kernel_radius);
std::vector<float> build_ring_kernel(const hp_kernel& k, const hp_ring* rings, uint32_t kernel_radius, int dim,
                                     uint32_t* size_out);
std::vector<float> folded_spectrum(const std::vector<float>& k, uint32_t size, int dim, uint32_t W, uint32_t H,
                                   uint32_t D);
hp_status seed_generate(const hp_seed* seed, uint32_t w, uint32_t h, uint32_t d, uint32_t c, float* const* out);

}  // namespace hp
