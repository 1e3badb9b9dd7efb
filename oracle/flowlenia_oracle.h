/*
 * flowlenia_oracle.h -- CPU oracle for the Flow Lenia time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * crate's CPU path (nathan-gage/autoverse, crate flow-lenia), used by the tests,
 * by the smoke check and as the CPU baseline leg of the benchmark.  The product
 * library (autoverse_amd) never links, loads or calls it.
 *
 * Every function is instantiated twice: suffix _f32 computes in float exactly as
 * the reference does (same operation order, no FMA contraction, glibc libm), and
 * suffix _f64 runs the same algorithm in double for error statements.
 *
 * Followed statement by statement (paths relative to the reference repository):
 *   kernel generation      src/compute/kernel.rs:32-117, src/compute/kernel3d.rs:32-134
 *   FFT convolution        src/compute/fft.rs:193-215, 288-393; src/compute/fft3d.rs:98-111, 308-455
 *                          (the FFT itself is this file's own mixed-radix Stockham code;
 *                          the reference uses rustfft 6.2, which is not in the tree)
 *   growth                 src/compute/growth.rs:11-59
 *   Sobel                  src/compute/gradient.rs:68-108, src/compute/gradient3d.rs:41-143
 *   alpha and flow         src/compute/flow.rs:14-87, src/compute/flow3d.rs:50-73
 *   reintegration          src/compute/reintegration.rs:60-228, src/compute/reintegration3d.rs:56-198
 *   step orchestration     src/compute/propagator.rs:157-427, src/compute/propagator3d.rs:47-308
 *   seeds                  src/schema/seed.rs:160-524
 *   config validation      src/schema/config.rs:136-164
 *   centre of mass         src/compute/evolution/fitness.rs:196-218, 309-326
 *   adjoint (oracle_grad.inc)  no reference counterpart: the transpose of the steps above
 *   embedded parameters (oracle_embedded.inc)  src/compute/direct_convolution.rs:29-174,
 *                          src/compute/param_advection.rs:83-357, src/schema/embedding.rs:60-167,
 *                          src/compute/embedded_propagator.rs:161-362
 *
 * Parity pins (which check covers which function) are listed in the tests; the
 * only stage with no bit-level pin is the FFT (rustfft's exact rounding cannot
 * be reproduced without rustfft), which is pinned to tolerance by brute-force
 * direct convolution, numpy.fft in float64, and the reference's FFT identities.
 */
#ifndef FLOWLENIA_ORACLE_H
#define FLOWLENIA_ORACLE_H

#include <stdint.h>
#include "../include/flowlenia_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

int orc_wrap_coord(int32_t coord, uint32_t size);
int32_t orc_sat_i32_f32(float v);
int32_t orc_sat_i32_f64(double v);

#define ORC_DECLARE(T, S)                                                                      \
    uint32_t orc_kernel2d_##S(const hp_kernel* k, const hp_ring* rings, uint32_t kernel_radius, \
                              T* out);                                                         \
    uint32_t orc_kernel3d_##S(const hp_kernel* k, const hp_ring* rings, uint32_t kernel_radius, \
                              T* out);                                                         \
    void orc_pad2d_##S(const T* data, uint32_t size, uint32_t w, uint32_t h, T* out);          \
    void orc_pad3d_##S(const T* data, uint32_t size, uint32_t w, uint32_t h, uint32_t d, T* out); \
    void orc_fft_##S(T* data, uint32_t n, int inverse);                                        \
    void orc_fft2d_##S(T* data, uint32_t w, uint32_t h, int inverse);                          \
    void orc_fft3d_##S(T* data, uint32_t w, uint32_t h, uint32_t d, int inverse);              \
    T orc_growth_##S(T u, T mu, T sigma);                                                       \
    void orc_growth_accumulate_##S(const T* conv, T* target, uint64_t n, T weight, T mu, T sigma); \
    void orc_sobel2d_##S(const T* grid, T* gx, T* gy, uint32_t w, uint32_t h);                 \
    void orc_sobel3d_##S(const T* grid, T* gx, T* gy, T* gz, uint32_t w, uint32_t h, uint32_t d); \
    T orc_alpha_##S(T mass, T beta_a, T n);                                                     \
    void orc_flow2d_##S(const T* gux, const T* guy, const T* gax, const T* gay, const T* mass,  \
                        uint64_t n, T beta_a, T nexp, T* fx, T* fy);                           \
    void orc_flow3d_##S(const T* gux, const T* guy, const T* guz, const T* gax, const T* gay,   \
                        const T* gaz, const T* mass, uint64_t n, T beta_a, T nexp, T* fx, T* fy, \
                        T* fz);                                                                \
    void orc_advect2d_##S(const T* current, const T* fx, const T* fy, T* next, uint32_t w,      \
                          uint32_t h, T dt, T s);                                              \
    void orc_advect3d_##S(const T* current, const T* fx, const T* fy, const T* fz, T* next,     \
                          uint32_t w, uint32_t h, uint32_t d, T dt, T s);                      \
    void orc_convolve_##S(const T* input, const T* kernel_freq, uint32_t w, uint32_t h,        \
                          uint32_t d, T* output);                                              \
    typedef struct orc_prop_##S orc_prop_##S;                                                  \
    orc_prop_##S* orc_prop_create_##S(const hp_config* cfg, int threads);                      \
    void orc_prop_destroy_##S(orc_prop_##S* p);                                                \
    void orc_prop_set_threads_##S(orc_prop_##S* p, int threads);                               \
    void orc_prop_set_line_threads_##S(orc_prop_##S* p, int threads);                          \
    void orc_prop_set_conserve_##S(orc_prop_##S* p, int conserve);                             \
    void orc_advect2d_mode_##S(const T* current, const T* fx, const T* fy, T* next, uint32_t w, \
                               uint32_t h, T dt, T s, int conserve);                           \
    void orc_advect3d_mode_##S(const T* current, const T* fx, const T* fy, const T* fz, T* next,\
                               uint32_t w, uint32_t h, uint32_t d, T dt, T s, int conserve);   \
    const T* orc_prop_kernel_freq_##S(const orc_prop_##S* p, uint32_t k);                      \
    void orc_prop_affinity_##S(orc_prop_##S* p, T* const* channels, T* const* affinity);        \
    void orc_prop_stencil_##S(orc_prop_##S* p, T* const* channels, T* const* affinity,          \
                              T* const* next);                                                 \
    void orc_prop_step_##S(orc_prop_##S* p, T* const* channels, float* time, uint64_t* step);   \
    void orc_prop_run_##S(orc_prop_##S* p, T* const* channels, float* time, uint64_t* step,     \
                          uint64_t steps);                                                     \
    orc_prop_##S* orc_prop_create_theta_##S(const hp_config* cfg, int threads, const double* theta); \
    uint32_t orc_theta_count_##S(const orc_prop_##S* p);                                       \
    int orc_grad_##S(orc_prop_##S* p, T* const* A0, uint64_t steps, T* const* gAT, T* const* AT, \
                     T* const* gA0, double* dtheta);                                            \
    /* embedded-parameter step (oracle_embedded.inc); params: 5 T per cell */                  \
    void orc_conv_direct2d_##S(const T* in, const T* k, uint32_t ksize, uint32_t w, uint32_t h, T* out); \
    void orc_conv_growth_emb_##S(const T* in, const T* k, uint32_t ksize, const T* params, uint32_t w, \
                                 uint32_t h, T* target, int accumulate, int threads);           \
    void orc_flow_emb_##S(const T* gux, const T* guy, const T* gax, const T* gay, const T* mass, \
                          const T* params, uint64_t n, T* fx, T* fy);                           \
    T orc_overlap_emb_##S(T dest_x, T dest_y, T s, T cell_x, T cell_y);                        \
    T orc_max_flow_mag_##S(const T* fx, const T* fy, uint64_t n);                              \
    void orc_mix_softmax_##S(const T* params, const T* mass, uint32_t count, T temperature, T* out); \
    void orc_mix_linear_##S(const T* params, const T* mass, uint32_t count, T* out);           \
    int32_t orc_advect_params_##S(const T* mass, const T* params, const T* fx, const T* fy, int linear, \
                                  T temperature, T dt, T s, uint32_t w, uint32_t h, T* next_mass, \
                                  T* next_params);                                              \
    typedef struct orc_emb_##S orc_emb_##S;                                                    \
    orc_emb_##S* orc_emb_create_##S(const hp_config* cfg, const hp_embedding* emb, int threads); \
    void orc_emb_destroy_##S(orc_emb_##S* p);                                                  \
    T* orc_emb_next_params_##S(orc_emb_##S* p, uint32_t c);                                    \
    int32_t orc_emb_last_radius_##S(const orc_emb_##S* p, uint32_t c);                         \
    void orc_emb_affinity_##S(orc_emb_##S* p, T* const* channels, T* const* params, T* const* affinity); \
    void orc_emb_stencil_##S(orc_emb_##S* p, T* const* channels, T* const* params, T* const* affinity, \
                             T* const* next, T* const* next_params);                            \
    void orc_emb_step_##S(orc_emb_##S* p, T* const* channels, T* const* params, float* time, uint64_t* step); \
    void orc_emb_run_##S(orc_emb_##S* p, T* const* channels, T* const* params, float* time,      \
                         uint64_t* step, uint64_t steps);

ORC_DECLARE(float, f32)
ORC_DECLARE(double, f64)

/* SimulationConfig::validate (config.rs:136-164). */
int orc_validate(const hp_config* cfg, uint32_t* bad_kernel, uint32_t* bad_channel);
/* Seed::generate (seed.rs:160-304), f32 as the reference. */
int orc_seed_generate(const hp_seed* seed, uint32_t w, uint32_t h, uint32_t d, uint32_t c,
                      float* const* out);
/* total_mass_all_channels (reintegration.rs:221-228): f32, sequential. */
float orc_total_mass_f32(const float* const* channels, uint32_t c, uint64_t n);
/* compute_center_of_mass (fitness.rs:196-218). */
void orc_center_of_mass(const float* grid, uint32_t w, uint32_t h, float* cx, float* cy);
/* compute_locomotion (fitness.rs:309-326) for a two-sample trajectory. */
float orc_locomotion(float x0, float y0, float x1, float y1, uint32_t w, uint32_t h);

#ifdef __cplusplus
}
#endif
#endif
