/*
 * flowlenia_hip.h -- C ABI of the MI355X-native Flow Lenia time step.
 *
 * The library is a drop-in for the reference crate's CPU propagator: same
 * configuration, same seed patterns, same state layout, same step semantics.
 * Every entry point is extern "C", takes plain pointers and sizes, never throws
 * and returns an hp_status.  Host buffers are owned by the caller and are never
 * retained; a context owns all device memory it allocates.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   hp_config / hp_kernel / hp_ring / hp_flow  <- SimulationConfig, KernelConfig,
 *       RingConfig, FlowConfig           (src/schema/config.rs:14-120)
 *   hp_validate_config                      <- SimulationConfig::validate
 *                                              (src/schema/config.rs:136-164)
 *   hp_seed / hp_seed_generate              <- Seed / Pattern / Seed::generate
 *                                              (src/schema/seed.rs:13-304)
 *   hp_create                               <- CpuPropagator::new
 *                                              (src/compute/propagator.rs:157-236)
 *                                              and CpuPropagator3D::new
 *                                              (src/compute/propagator3d.rs:49-117)
 *   hp_step                                 <- CpuPropagator::step
 *                                              (src/compute/propagator.rs:239-420)
 *   hp_run                                  <- CpuPropagator::run (propagator.rs:423-427)
 *   hp_get_config                           <- CpuPropagator::config (propagator.rs:430-432)
 *   hp_mass_f64                             <- total_mass_all_channels
 *                                              (src/compute/reintegration.rs:221-228),
 *                                              computed in FP64 instead of f32
 *   hp_advect_mass_2d / hp_advect_mass_3d / hp_sobel_gradient_2d /
 *   hp_compute_flow_2d / hp_growth_accumulate
 *                                           <- the stage functions the crate
 *       re-exports from src/compute/mod.rs:5-41 (advect_mass_into
 *       reintegration.rs:60-88, advect_mass_3d_into reintegration3d.rs:56-101,
 *       sobel_gradient_into gradient.rs:68-108, compute_flow_field_into
 *       flow.rs:68-87, growth_accumulate growth.rs:44-59)
 *   hp_affinity                             <- stage 1 of CpuPropagator::step
 *       (propagator.rs:244-280): every kernel's
 *       CachedConvolver::convolve_with_kernel_scratch (fft.rs:288-324) plus
 *       growth_accumulate, on a context's resident kernel spectra; there is
 *       no single-convolution entry point
 *   hp_stencil_with_affinity                <- stages 2-5 of CpuPropagator::step
 *       (propagator.rs:303-414) with U supplied by the caller
 */
#ifndef FLOWLENIA_HIP_H
#define FLOWLENIA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The first five mirror ConfigError (src/schema/config.rs:169-180). */
typedef enum hp_status {
    HP_OK = 0,
    HP_ERR_INVALID_DIMENSIONS = 1,    /* width, height or depth is zero            */
    HP_ERR_INVALID_CHANNELS = 2,      /* channels is zero                          */
    HP_ERR_INVALID_TIME_STEP = 3,     /* dt <= 0                                   */
    HP_ERR_INVALID_KERNEL_RADIUS = 4, /* kernel_radius is zero                     */
    HP_ERR_INVALID_CHANNEL_INDEX = 5, /* a kernel names a channel >= channels      */
    HP_ERR_NO_DEVICE = 6,             /* no HIP device, or the ordinal is invalid  */
    HP_ERR_HIP = 7,                   /* a HIP runtime call failed                 */
    HP_ERR_SHAPE = 8,                 /* buffer count / size mismatch              */
    HP_ERR_NULL = 9,                  /* a required pointer is NULL                */
    HP_ERR_NONFINITE = 10,            /* the state holds NaN or Inf                */
    HP_ERR_UNSUPPORTED_SIZE = 11,     /* a size outside what the kernels support   */
    HP_ERR_IO = 12,                   /* a file could not be opened, read or written,
                                         or is not a valid FLWA animation            */
    HP_ERR_PARSE = 13                 /* JSON text that is not a valid config / seed */
} hp_status;

/* RingConfig (config.rs:92-99). */
typedef struct hp_ring {
    float amplitude;
    float distance;
    float width;
} hp_ring;

/* KernelConfig (config.rs:55-70).  The rings of kernel i are
 * rings[ring_offset .. ring_offset + ring_count) of hp_config.rings. */
typedef struct hp_kernel {
    float radius;
    float weight;
    float mu;
    float sigma;
    uint32_t source_channel;
    uint32_t target_channel;
    uint32_t ring_offset;
    uint32_t ring_count;
} hp_kernel;

/* FlowConfig (config.rs:103-110). */
typedef struct hp_flow {
    float beta_a;
    float n;
    float distribution_size;
} hp_flow;

/* SimulationConfig (config.rs:14-35).  depth == 1 selects the 2D step. */
typedef struct hp_config {
    uint32_t width;
    uint32_t height;
    uint32_t depth;
    uint32_t channels;
    float dt;
    uint32_t kernel_radius;
    hp_flow flow;
    uint32_t num_kernels;
    const hp_kernel* kernels;
    uint32_t num_rings;
    const hp_ring* rings;
} hp_config;

/* SimulationConfig::validate (config.rs:136-164).  On HP_ERR_INVALID_CHANNEL_INDEX,
 * *bad_kernel / *bad_channel (if non-NULL) receive the offending pair. */
hp_status hp_validate_config(const hp_config* cfg, uint32_t* bad_kernel, uint32_t* bad_channel);

/* ---- Seeds (src/schema/seed.rs) ------------------------------------------ */
typedef enum hp_pattern_kind {
    HP_PATTERN_GAUSSIAN_BLOB = 0,   /* seed.rs:40-49   */
    HP_PATTERN_MULTI_BLOB = 1,      /* seed.rs:51-54   */
    HP_PATTERN_RING = 2,            /* seed.rs:56-68   */
    HP_PATTERN_GAUSSIAN_SPHERE = 3, /* seed.rs:71-80   */
    HP_PATTERN_MULTI_SPHERE = 4,    /* seed.rs:82-85   */
    HP_PATTERN_SHELL = 5,           /* seed.rs:87-98   */
    HP_PATTERN_TORUS3D = 6,         /* seed.rs:100-111 */
    HP_PATTERN_NOISE = 7,           /* seed.rs:115-122 */
    HP_PATTERN_CUSTOM = 8,          /* seed.rs:125-128 */
    HP_PATTERN_CUSTOM3D = 9         /* seed.rs:130-134 */
} hp_pattern_kind;

/* One blob (BlobSpec, seed.rs:138-143) or sphere (SphereSpec, seed.rs:146-152).
 * For blobs center[2] is ignored. */
typedef struct hp_blob {
    float center[3];
    float radius;
    float amplitude;
    uint32_t channel;
} hp_blob;

/* One Custom / Custom3D entry; z is ignored for Custom. */
typedef struct hp_custom_value {
    uint32_t x, y, z, channel;
    float value;
} hp_custom_value;

typedef struct hp_seed {
    int32_t kind;                 /* hp_pattern_kind */
    float center[3];              /* GaussianBlob/Ring use center[0..1] */
    float radius;                 /* GaussianBlob, GaussianSphere */
    float inner_radius;           /* Ring, Shell */
    float outer_radius;           /* Ring, Shell */
    float major_radius;           /* Torus3D */
    float minor_radius;           /* Torus3D */
    float amplitude;
    int32_t channel;              /* Noise: -1 means every channel */
    uint64_t noise_seed;          /* Noise */
    uint32_t num_blobs;           /* MultiBlob / MultiSphere */
    const hp_blob* blobs;
    uint32_t num_values;          /* Custom / Custom3D */
    const hp_custom_value* values;
} hp_seed;

/* Seed::generate (seed.rs:160-304): rasterise into channels[c][z*H*W + y*W + x].
 * The caller passes `channels` host buffers of width*height*depth floats; they
 * are overwritten (zeroed first).  Pure host code, no device needed. */
hp_status hp_seed_generate(const hp_seed* seed, uint32_t width, uint32_t height, uint32_t depth,
                           uint32_t channels, float* const* out_channels);

/* ---- JSON files (serde_json's reading of the reference's schema) ------------------------
 * hp_config_from_json / hp_seed_from_json replace serde_json::from_str::<SimulationConfig> and
 * ::<Seed> (config.rs:14-120, seed.rs:13-152): the same required fields and defaults (depth 1,
 * embedding disabled), the "type"-tagged Pattern, tuples as arrays of their exact length,
 * integers where the reference has usize / u64, unknown fields ignored.  The documents own the
 * arrays their hp_config / hp_seed point into.  Failures return HP_ERR_PARSE with a message
 * naming the field (hp_json_last_error). */
typedef struct hp_config_doc hp_config_doc;
typedef struct hp_seed_doc hp_seed_doc;

/* EmbeddingConfig (src/schema/embedding.rs:326-345). */
typedef struct hp_embedding {
    int32_t enabled;
    float mixing_temperature;
    int32_t linear_mixing;
} hp_embedding;

hp_status hp_config_from_json(const char* text, hp_config_doc** out);
hp_status hp_seed_from_json(const char* text, hp_seed_doc** out);
const hp_config* hp_config_doc_config(const hp_config_doc* doc);
hp_status hp_config_doc_embedding(const hp_config_doc* doc, hp_embedding* out);
void hp_config_doc_free(hp_config_doc* doc);
const hp_seed* hp_seed_doc_seed(const hp_seed_doc* doc);
void hp_seed_doc_free(hp_seed_doc* doc);
/* load_config_and_seed (main.rs:409-436): the config from config_path; the seed from the same
 * path with its extension replaced by "seed.json", or Seed::default() (a GaussianBlob at
 * (0.5, 0.5), radius 0.1, amplitude 1, channel 0) when that file does not exist.  HP_ERR_IO if a
 * file cannot be read. */
hp_status hp_load_config_and_seed(const char* config_path, hp_config_doc** config, hp_seed_doc** seed);
const char* hp_json_last_error(void);

/* ---- Propagator ----------------------------------------------------------- */
typedef struct hp_ctx hp_ctx;

/* Build the propagator: validates the config, builds the ring kernels
 * (kernel.rs:32-86 / kernel3d.rs:32-97), their spectra, and every device buffer.
 * `device` is a HIP device ordinal.  Returns HP_ERR_NO_DEVICE when there is no
 * GPU; *out is NULL on any error.  Any side is accepted: sides 2^a * 3^b from 8 to 4096
 * convolve by FFT, others by a direct tap sum (the same circular convolution to rounding;
 * kernel radius up to 85 in 2D, 13 in 3D).  HP_ERR_UNSUPPORTED_SIZE when a kernel's
 * diameter 2R+1 exceeds a side (the reference panics). */
hp_status hp_create(const hp_config* cfg, int device, hp_ctx** out);
void hp_destroy(hp_ctx* ctx);

/* Message for the last error on this context (valid until the next call). */
const char* hp_last_error(const hp_ctx* ctx);
/* Message for the last error of a call that had no context (hp_create). */
const char* hp_last_global_error(void);

/* Copy of the config the context was built from; kernels/rings point into the
 * context and stay valid until hp_destroy. */
hp_status hp_get_config(const hp_ctx* ctx, hp_config* out);

/* Upload a state: `channels` host pointers, width*height*depth floats each. */
hp_status hp_set_state(hp_ctx* ctx, const float* const* channels, float time, uint64_t step);
/* Download the device-resident state. */
hp_status hp_get_state(hp_ctx* ctx, float* const* channels, float* time, uint64_t* step);

/* CpuPropagator::step semantics on a host state: upload, one step, download,
 * time += dt (f32), step += 1. */
hp_status hp_step(hp_ctx* ctx, float* const* channels, float* time, uint64_t* step);
/* CpuPropagator::run semantics: one upload, `steps` device-resident steps, one download. */
hp_status hp_run(hp_ctx* ctx, float* const* channels, float* time, uint64_t* step, uint64_t steps);
/* Advance the device-resident state by `steps` steps with no host transfer
 * (the state set by hp_set_state or left by earlier calls).  Steps are replayed
 * from a hipGraph; synchronises before returning. */
hp_status hp_advance(hp_ctx* ctx, uint64_t steps);

/* FP64 per-channel and total mass of the device-resident state, reduced in a
 * fixed order (bitwise reproducible run to run).  per_channel may be NULL. */
hp_status hp_mass_f64(hp_ctx* ctx, double* per_channel, double* total);

/* Reintegration arithmetic (no reference counterpart).  HP_CONSERVATION_REFERENCE (the default)
 * reproduces reintegration.rs's f32 arithmetic bit for bit; its total mass drifts (about 2e-6
 * relative per 1000 steps) because box widths in absolute coordinates are not exactly 2s where a
 * box straddles a power of two and because every destination sum rounds at each addition.
 * HP_CONSERVATION_EXACT scales each source's mass once to q = m / (product of its overlap-width
 * sums), forms the contributions q * (ow * oh [* od]) and every destination's sum in FP64 in the
 * scatter's order, and rounds once: the drift falls to the FP64 reference's level, at a parity
 * cost after 1000 steps of 2e-5 max-abs against the reference arithmetic on the plain 256^2
 * configuration (4e-4 on the 1024^2 x 3 multispecies one, whose dynamics amplify differences).
 * Changing the mode keeps the state; the step's replay graphs are captured again. */
enum { HP_CONSERVATION_REFERENCE = 0, HP_CONSERVATION_EXACT = 1 };
hp_status hp_set_conservation(hp_ctx* ctx, int mode);
hp_status hp_get_conservation(const hp_ctx* ctx, int* mode);

/* Largest gather radius any step has needed so far, and how many steps had to
 * take the general reintegration path (a flow too large for the fused stencil). */
hp_status hp_stats(const hp_ctx* ctx, uint32_t* max_gather_radius, uint64_t* general_path_steps);

/* ---- Execution plan and its autotuner ------------------------------------------
 * No counterpart in the reference (CpuPropagator has one code path).  A plan picks how
 * the step runs on the device; every plan computes the same step: tile shapes and spectral
 * grouping change nothing bit for bit, FFT variants only the FFT's rounding.
 * hp_create uses the plan the autotuner last found for a configuration of the same shape in
 * this process, else built-in defaults. */
typedef struct hp_plan {
    int32_t fft_path;       /* 0: wave-level column-major FFTs where the sides allow (2D, 256..2048),
                               1: LDS FFTs (always available) */
    int32_t spectral_group; /* 1: one spectral work item per source channel, 0: one per kernel,
                               -1: chosen from the work size */
    int32_t radix_order;    /* LDS FFTs: 0 radix 8 first, 1 radix 4 first */
    int32_t column_split;   /* LDS spectral pass: 1 or 2 (half the columns per workgroup) */
    int32_t stencil_tile;   /* fused 2D stencil tile (destinations / threads): 0 64x16/512, 1 64x8/256,
                               2 32x16/256, 3 128x8/512, 4 64x16/256, 5 128x8/256 */
} hp_plan;

/* The context's resolved plan (fft_path and spectral_group as actually used). */
hp_status hp_get_plan(const hp_ctx* ctx, hp_plan* out);
/* Rebuild the device plan; the state, time and step are kept. */
hp_status hp_set_plan(hp_ctx* ctx, const hp_plan* plan);
/* Time every candidate plan over `steps` steps from the current state (which is restored
 * afterwards, with time and step), keep the fastest whose result agrees with the default
 * plan's to 1e-3, and remember it for later contexts of the same shape.  Up to max_results
 * candidates are reported in tried / us_per_step (either may be NULL); *n_tried (optional)
 * receives how many were timed. */
hp_status hp_autotune(hp_ctx* ctx, uint32_t steps, hp_plan* best, uint32_t max_results, hp_plan* tried,
                      double* us_per_step, uint32_t* n_tried);

/* ---- Per-kernel timing (HIP events on the context's stream) ------------------ */
/* Kernel classes of the step, in launch order. */
typedef enum hp_kernel_class {
    HP_K_ROWS_FWD = 0,   /* R2C along x (2D rows / 3D x-lines)                 */
    HP_K_COLS_FWD = 1,   /* forward FFT along y (3D only)                        */
    HP_K_SPECTRAL = 2,   /* last forward axis, x kernel spectrum, inverse axis */
    HP_K_COLS_INV = 3,   /* inverse FFT along y (3D only)                        */
    HP_K_ROWS_INV = 4,   /* C2R along x + growth + weighted accumulate           */
    HP_K_STENCIL = 5,    /* Sobel + flow + gather reintegration                  */
    HP_K_GENERAL = 6,    /* general-radius reintegration (rare path)             */
    HP_K_NUM_CLASSES = 7
} hp_kernel_class;

/* Turn per-kernel event timing on or off (off by default; on adds one event pair
 * per launch and disables graph replay).  Enabling resets the counters. */
hp_status hp_profile_enable(hp_ctx* ctx, int enable);
/* Profiling aid, no reference counterpart: in profiling mode, launch each pass that writes
 * only buffers its inputs do not alias (the wave-path FFT passes, the fused stencils) `reps`
 * times back to back inside its event pair, so that the per-launch average carries 1/reps of
 * the event overhead.  1..1000; default 1. */
hp_status hp_profile_repeats(hp_ctx* ctx, uint32_t reps);
/* Total milliseconds and launch count per class since enabling. */
hp_status hp_profile_read(hp_ctx* ctx, double* ms /*HP_K_NUM_CLASSES*/, uint64_t* launches /*HP_K_NUM_CLASSES*/);
/* Algorithmic HBM bytes one launch of each class moves for this config. */
hp_status hp_kernel_bytes(const hp_ctx* ctx, double* bytes /*HP_K_NUM_CLASSES*/);

/* ---- Stage functions (host buffers in and out, run on `device`) -------------- */
/* advect_mass_into (reintegration.rs:60-88): next = reintegrate(current, flow);
 * `next` is overwritten (the reference adds into a zeroed buffer). */
hp_status hp_advect_mass_2d(int device, const float* current, const float* flow_x, const float* flow_y,
                            float* next, uint32_t width, uint32_t height, float dt, float distribution_size);
/* advect_mass_3d_into (reintegration3d.rs:56-101). */
hp_status hp_advect_mass_3d(int device, const float* current, const float* flow_x, const float* flow_y,
                            const float* flow_z, float* next, uint32_t width, uint32_t height,
                            uint32_t depth, float dt, float distribution_size);
/* No counterpart in the reference.  hp_advect_mass_2d for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * `current`, `flow_x`, `flow_y` and `next` are the caller's DEVICE buffers of width*height
 * floats on `device`; the result is hp_advect_mass_2d's bit for bit.  The per-source boxes and
 * flags live in a static device scratch of 32 MiB per GPU, so grids above 1,048,576 cells are
 * refused with HP_ERR_UNSUPPORTED_SIZE.  Calls that overlap on one GPU share that scratch and
 * race.  Runs on the null stream and synchronises before returning. */
hp_status hp_advect_mass_2d_noalloc(int device, const float* current, const float* flow_x, const float* flow_y,
                                    float* next, uint32_t width, uint32_t height, float dt,
                                    float distribution_size);
/* sobel_gradient_into (gradient.rs:68-108). */
hp_status hp_sobel_gradient_2d(int device, const float* grid, float* grad_x, float* grad_y,
                               uint32_t width, uint32_t height);
/* compute_flow_field_into (flow.rs:68-87). */
hp_status hp_compute_flow_2d(int device, const float* grad_u_x, const float* grad_u_y,
                             const float* grad_a_x, const float* grad_a_y, const float* mass_sum,
                             uint32_t n_cells, hp_flow flow, float* flow_x, float* flow_y);
/* growth_accumulate (growth.rs:44-59): target[i] += weight * G(conv[i]). */
hp_status hp_growth_accumulate(int device, const float* conv, float* target, uint32_t n_cells,
                               float weight, float mu, float sigma);
/* Flow-field reintegration of a context's step fed with a given affinity:
 * next = stencil(state, affinity) for every channel, i.e. stages 2-5 of
 * CpuPropagator::step (propagator.rs:303-414) with U supplied by the caller. */
hp_status hp_stencil_with_affinity(hp_ctx* ctx, const float* const* state, const float* const* affinity,
                                   float* const* next);
/* Affinity U of the context's step for a given state: stage 1 of
 * CpuPropagator::step (propagator.rs:244-280): FFT convolution + growth. */
hp_status hp_affinity(hp_ctx* ctx, const float* const* state, float* const* affinity);

/* ---- Populations ------------------------------------------------------------- */
/* A population of independent simulations, possibly of different sizes and channel
 * counts, stepped together: every pass of a step is one launch per size class (2D),
 * all passes of a step replay from one hipGraph.  Replaces the per-candidate
 * `CpuPropagator::new` + `step` loop of FitnessEvaluator::evaluate
 * (src/compute/evolution/fitness.rs:26-54) driven by EvolutionEngine::evaluate_population
 * (src/compute/evolution/search.rs:153-172).  2D members only. */
typedef struct hp_batch hp_batch;
hp_status hp_batch_create(const hp_config* members, uint32_t count, int device, hp_batch** out);
void hp_batch_destroy(hp_batch* b);
const char* hp_batch_last_error(const hp_batch* b);
/* Upload / download one member's state (its `channels` host buffers, W*H floats each). */
hp_status hp_batch_set_state(hp_batch* b, uint32_t member, const float* const* channels);
hp_status hp_batch_get_state(hp_batch* b, uint32_t member, float* const* channels, float* time, uint64_t* step);
/* Advance every member by `steps` steps, device-resident; synchronises before returning. */
hp_status hp_batch_advance(hp_batch* b, uint64_t steps);
/* Centre of mass of every member's channel sum (compute_center_of_mass, fitness.rs:196-218,
 * accumulated in FP64): out[2*i], out[2*i+1] = (cx, cy); the grid centre when the mass is
 * <= 1e-6.  mass (may be NULL) receives the FP64 positive mass. */
hp_status hp_batch_centroids(hp_batch* b, double* out_xy, double* mass);
/* FP64 total mass of every member (fixed reduction order). */
hp_status hp_batch_mass_f64(hp_batch* b, double* per_member);
hp_status hp_batch_stats(const hp_batch* b, uint32_t* max_gather_radius, uint64_t* general_path_steps);

/* ---- fitness evaluation (src/compute/evolution/fitness.rs) ------------------------------
 * hp_batch_evaluate replaces FitnessEvaluator::evaluate (fitness.rs:26-86) for every member at
 * once: from each member's current state it records the trajectory's initial mass and
 * snapshot, runs the warmup, samples at step 0 and every sample_interval steps
 * (EvaluationTrajectory::record_sample, fitness.rs:146-160), runs the remaining steps, then
 * computes each metric (compute_metric, fitness.rs:247-268) and the weighted combination.  The
 * sums behind every sample (mass, centre of mass, radius, peak, active cells, similarity dot
 * products) are FP64 reductions on the device; the metric formulas are the reference's, in f32,
 * on the host.  The members are left holding their final states. */
/* EvaluationConfig (src/schema/evolution.rs:479-495); sample_interval 0 is taken as 1. */
typedef struct hp_eval_config {
    uint64_t steps;
    uint64_t warmup_steps;
    uint64_t sample_interval;
} hp_eval_config;

/* FitnessMetric (src/schema/evolution.rs:319-375). */
typedef enum hp_metric_kind {
    HP_METRIC_PERSISTENCE = 0,
    HP_METRIC_COMPACTNESS = 1,
    HP_METRIC_LOCOMOTION = 2,
    HP_METRIC_PERIODICITY = 3,        /* period, tolerance */
    HP_METRIC_COMPLEXITY = 4,
    HP_METRIC_MASS_CONCENTRATION = 5,
    HP_METRIC_GLIDER_SCORE = 6,       /* min_displacement */
    HP_METRIC_OSCILLATOR_SCORE = 7,   /* max_period, threshold */
    HP_METRIC_STABILITY = 8,
    HP_METRIC_CUSTOM = 9              /* scored outside the evaluator: 0 */
} hp_metric_kind;

/* WeightedMetric (src/schema/evolution.rs:309-314) with the variant's parameters. */
typedef struct hp_metric {
    int32_t kind;
    float weight;
    uint64_t period;
    float tolerance;
    float min_displacement;
    uint64_t max_period;
    float threshold;
} hp_metric;

/* BehaviorStats (src/schema/evolution.rs:1015-1032) without its two per-sample series, which
 * hp_sample carries. */
typedef struct hp_behavior {
    float final_mass;
    float initial_mass;
    float total_displacement;
    float final_radius;
    uint64_t active_cells;
    float max_activation;
} hp_behavior;

/* One recorded sample of a member's channel sum. */
typedef struct hp_sample {
    float center_x;
    float center_y;
    float radius;
    float mass;
    float max_value;
    uint32_t active_cells;
} hp_sample;

/* Samples one evaluation records: 1 + steps / max(sample_interval, 1). */
hp_status hp_eval_sample_count(const hp_eval_config* ev, uint32_t* count);
/* fitness[M] receives the combined fitness (FitnessConfig::normalize clamps each score to
 * [0, 1] before weighting), scores[M * num_metrics] the raw metric scores; behavior[M] and
 * samples[M * count] may be NULL. */
hp_status hp_batch_evaluate(hp_batch* b, const hp_eval_config* ev, const hp_metric* metrics, uint32_t num_metrics,
                            int normalize, float* fitness, float* scores, hp_behavior* behavior,
                            hp_sample* samples);

/* ---- gradient tape (2D members) ---------------------------------------------------------
 * No counterpart in the reference, which has no derivative.  A tape records T steps of the
 * whole population from its current state (the state advances to A_T, as hp_batch_advance),
 * keeping one state every `checkpoint_every` steps; a backward pass then gives, for the loss
 * L = sum_c <dl_dfinal_c, A_T,c> of every member, dL/dA_0 and dL/dtheta, recomputing each
 * segment of steps from its checkpoint.  theta of a member is, per kernel in config order,
 * (weight, mu, sigma), then per ring in config order (amplitude, distance, width).
 * Pointer arrays are member-major: member 0's channels, then member 1's, ...; dtheta holds
 * the members' theta blocks back to back.  The tape is dropped by set_state / advance.
 * checkpoint_every = 0 picks the interval from free device memory: every step (no
 * recomputation) when the whole trajectory fits in half of it, else about sqrt(steps);
 * *used_every (optional) receives the interval used.
 * Limits: sides 2^a * 3^b (8 .. 4096); every kernel's support 2R+1 within the grid. */
hp_status hp_batch_theta_count(const hp_batch* b, uint32_t member, uint32_t* count);
/* Batch plans: the LDS FFT variant and the stencil tile shared by all members (fft_path and
 * spectral_group are fixed for populations: LDS FFTs, one spectral item per source).
 * hp_batch_autotune times the candidates from the current states and restores them. */
hp_status hp_batch_get_plan(const hp_batch* b, hp_plan* out);
hp_status hp_batch_set_plan(hp_batch* b, const hp_plan* plan);
hp_status hp_batch_autotune(hp_batch* b, uint32_t steps, hp_plan* best, uint32_t max_results, hp_plan* tried,
                            double* us_per_step, uint32_t* n_tried);
/* Step mode of a population (no reference counterpart: the reference steps each candidate's
 * CpuPropagator on its own, fitness.rs:26-54).  HP_STEP_MULTI_LAUNCH (the default): a step is one
 * launch per FFT pass and size class and one stencil launch per channel-count group, replayed
 * from hipGraphs.  HP_STEP_PERSISTENT: a step is one launch, replayed from a hipGraph, whose
 * workgroups (as many as the device holds at once) walk every member's forward row pass, then
 * every spectral pass, inverse row pass and stencil tile, with a bounded grid barrier between
 * phases; every member's sides must be 64, 128, 192 or 256 (else HP_ERR_UNSUPPORTED_SIZE) and
 * the plan's radix order 0.  Both modes give every member the same trajectory bit for bit.  A
 * barrier that waits past its bound makes the run recompute its steps through the checked path
 * and counts one *barrier_timeouts. */
#define HP_STEP_MULTI_LAUNCH 0
#define HP_STEP_PERSISTENT 1
hp_status hp_batch_set_step_mode(hp_batch* b, int32_t mode);
hp_status hp_batch_step_mode(const hp_batch* b, int32_t* mode, uint64_t* barrier_timeouts);
hp_status hp_batch_tape_forward(hp_batch* b, uint64_t steps, uint32_t checkpoint_every, uint32_t* used_every);
hp_status hp_batch_tape_backward(hp_batch* b, const float* const* dl_dfinal, float* const* dl_dinitial,
                                 double* dtheta);

/* ---- Animation files (src/animation/): the FLWA format -----------------------------------
 * Layout (format.rs:60-185), little-endian: a 52-byte header -- "FLWA", version u16 = 1, flags
 * u16 (compression in bits 0-3: 0 none, 1 LZ4 with a 4-byte size prefix; bit 4 delta encoding),
 * width, height, depth, channels (u32 each), frame_count u64, dt f32, 16 zero bytes -- then the
 * frames, each the state's channels back to back as f32 ([C][D*H*W], the device layout of a
 * context's state), then one 16-byte index entry (offset u64, size u64) per frame. */
typedef enum hp_compression { HP_COMPRESSION_NONE = 0, HP_COMPRESSION_LZ4 = 1 } hp_compression;

/* RecorderConfig (recorder.rs:14-33): record every frame_skip-th call (0 taken as 1), at most
 * max_frames frames (0 = no limit). */
typedef struct hp_recorder_config {
    int32_t compression;
    uint32_t frame_skip;
    uint64_t max_frames;
} hp_recorder_config;

/* AnimationStats (recorder.rs:214-224). */
typedef struct hp_animation_stats {
    uint64_t frame_count;
    uint64_t total_bytes;
    uint64_t average_frame_size;
    int32_t compression;
} hp_animation_stats;

/* The header fields (AnimationHeader, format.rs:60-76). */
typedef struct hp_animation_info {
    uint32_t width, height, depth, channels;
    uint64_t frame_count;
    float dt;
    int32_t compression;
    int32_t delta_encoding;
} hp_animation_info;

typedef struct hp_recorder hp_recorder;
/* AnimationRecorder::new (recorder.rs:63-114): create `path` and write the placeholder header
 * for a configuration's grid. */
hp_status hp_recorder_open(const char* path, const hp_config* cfg, const hp_recorder_config* rc, hp_recorder** out);
/* AnimationRecorder::record_frame (recorder.rs:119-170).  The _ctx and _member forms take the
 * frame straight from device memory: one device-to-host copy into a pinned staging buffer,
 * queued on the context's (population's) stream behind its steps, and written to the file by
 * a background thread while the device goes on stepping.  *recorded (may be NULL) is 1 when
 * this call produced a frame (frame_skip / max_frames). */
hp_status hp_recorder_record_ctx(hp_recorder* r, hp_ctx* ctx, int* recorded);
hp_status hp_recorder_record_member(hp_recorder* r, hp_batch* b, uint32_t member, int* recorded);
/* The same from host buffers (one pointer per channel, W*H*D floats each). */
hp_status hp_recorder_record_host(hp_recorder* r, const float* const* channels, int* recorded);
/* AnimationRecorder::finalize (recorder.rs:175-205): drain the writer, write the index table,
 * rewrite the header with the frame count, close.  The recorder is freed (also on error). */
hp_status hp_recorder_finalize(hp_recorder* r, hp_animation_stats* stats);
/* Message of the last error of a recorder or player call on this thread. */
const char* hp_animation_last_error(void);

typedef struct hp_player hp_player;
/* AnimationPlayer::open (player.rs:36-64): header and index table. */
hp_status hp_player_open(const char* path, hp_player** out);
hp_status hp_player_info(const hp_player* p, hp_animation_info* out);
/* AnimationPlayer::read_frame_into (player.rs:130-180): frame `index` into `channels` host
 * buffers (one per channel, W*H*D floats each). */
hp_status hp_player_read_frame(hp_player* p, uint64_t index, float* const* channels);
/* Resume: frame `index` becomes the device-resident state of a context built for the same grid
 * and channel count, with the given time and step (hp_set_state semantics). */
hp_status hp_player_load_ctx(hp_player* p, uint64_t index, hp_ctx* ctx, float time, uint64_t step);
void hp_player_close(hp_player* p);

/* ---- embedded-parameter mode (src/compute/embedded_propagator.rs) -----------------------
 * The state carries per-cell parameters that move with the mass (EmbeddedState,
 * embedded_propagator.rs:24-37): per channel W*H CellParams (embedding.rs:20-31), five floats
 * each, laid out as the reference's Vec<CellParams>.  The grid is 2D; depth is not read.
 * Differences from the reference, all in the convolution stage: the convolutions run as FFTs
 * (the reference sums the kernel taps directly, direct_convolution.rs:129-174; the two agree
 * to the FFT's rounding), and exp is evaluated in FP64 and rounded once (libm's expf differs
 * from that by one ulp in about 7 of 10000 arguments).  Given the affinity, the flow, the
 * gather and the mixing follow the reference's f32 arithmetic operation for operation. */
typedef struct hp_cell_params {
    float mu, sigma, weight, beta_a, n;
} hp_cell_params;

typedef struct hp_emb hp_emb;
/* EmbeddedPropagator::new (embedded_propagator.rs:165-220); emb NULL is EmbeddingConfig's
 * default (disabled).  The propagator's next parameter grids start as CellParams::default(). */
hp_status hp_emb_create(const hp_config* cfg, const hp_embedding* emb, int device, hp_emb** out);
void hp_emb_destroy(hp_emb* e);
/* Message of the last failed call on e (e NULL: of the last failed hp_emb_create or
 * hp_advect_mass_and_params on this thread). */
const char* hp_emb_last_error(const hp_emb* e);
/* The parameters EmbeddedState::from_seed gives every cell (embedded_propagator.rs:50-57). */
hp_status hp_emb_default_params(const hp_config* cfg, hp_cell_params* out);
/* The device-resident state: channels (W*H floats each) and parameter grids (W*H cells each). */
hp_status hp_emb_set_state(hp_emb* e, const float* const* channels, const hp_cell_params* const* params, float time,
                           uint64_t step);
/* params may be NULL (mass only). */
hp_status hp_emb_get_state(hp_emb* e, float* const* channels, hp_cell_params* const* params, float* time,
                           uint64_t* step);
/* EmbeddedPropagator::step and ::run (embedded_propagator.rs:222-338) on host buffers, in
 * place.  A step whose search radius would exceed the gather's work bound (2e10 source visits)
 * returns HP_ERR_UNSUPPORTED_SIZE and leaves the state as it was. */
hp_status hp_emb_step(hp_emb* e, float* const* channels, hp_cell_params* const* params, float* time,
                      uint64_t* step);
hp_status hp_emb_run(hp_emb* e, float* const* channels, hp_cell_params* const* params, float* time, uint64_t* step,
                     uint64_t steps);
/* `steps` steps of the device-resident state. */
hp_status hp_emb_advance(hp_emb* e, uint64_t steps);
/* FP64 mass of the device-resident state (as hp_mass_f64). */
hp_status hp_emb_mass_f64(hp_emb* e, double* per_channel, double* total);
/* The propagator's next parameter grids: what a destination that receives no mix keeps. */
hp_status hp_emb_get_next_params(hp_emb* e, hp_cell_params* const* params);
hp_status hp_emb_set_next_params(hp_emb* e, const hp_cell_params* const* params);
/* Search radius of the last step per channel (param_advection.rs:99-100); 0 with embedding
 * disabled. */
hp_status hp_emb_last_radius(hp_emb* e, int32_t* radius);
/* Stage 1 (embedded_propagator.rs:229-249) for the given state: the affinity per channel. */
hp_status hp_emb_affinity(hp_emb* e, const float* const* channels, const hp_cell_params* const* params,
                          float* const* affinity);
/* Stages 2-5 (embedded_propagator.rs:251-330) with the affinity given; next_params is read and
 * written (destinations that receive no mix keep its values). */
hp_status hp_emb_stencil(hp_emb* e, const float* const* channels, const hp_cell_params* const* params,
                         const float* const* affinity, float* const* next, hp_cell_params* const* next_params);
/* advect_mass_and_params_into (param_advection.rs:83-166), one channel on host buffers;
 * next_params is read and written as above; emb NULL is EmbeddingConfig::enabled(). */
hp_status hp_advect_mass_and_params(const float* mass, const hp_cell_params* params, const float* fx,
                                    const float* fy, const hp_embedding* emb, float dt, float s, uint32_t w,
                                    uint32_t h, float* next_mass, hp_cell_params* next_params, int device);

#ifdef __cplusplus
}
#endif

#endif /* FLOWLENIA_HIP_H */
