/* include/sing_abi.h — C ABI of the MI355X offline null-geodesic renderer.
 *
 * Drop-in boundary for the reference's CUDA offline path:
 *   - render/include/render_backend.hpp:107-132   RenderBackend {initialize, shutdown,
 *     resize, render_frame, capture_frame, name}
 *   - render/cuda/cuda_backend.cu:27-28,144-148    singularity_geodesic_kernel(uint8_t*, Uniforms)
 *     launched by CudaBackend::render_frame
 *   - render/cuda/cuda_backend.cu:45-101           pack_uniforms_cuda(Scene, CameraState, w, h, t)
 *   - cuda_cli/main.cpp:68-109                     build_orbital_camera(...)
 *   - core/include/scene/scene_config.hpp:134-199  load_scene_config_from_{stream,file}
 *
 * Plain C types only: pointers, sizes, PODs.  Every entry point returns an int
 * status (SING_OK = 0) and never aborts; the per-context message is available
 * from sing_last_error().  Contexts own all device memory; host buffers passed
 * to capture calls belong to the caller.  One context per host thread; calls on
 * different contexts may run concurrently.
 */
#ifndef SING_ABI_H
#define SING_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SING_ABI_VERSION 3

/* --- the 160-byte uniform block, byte-for-byte shared_shader/uniforms.h:54-95 --- */
typedef struct SingVec4 {
    float x, y, z, w;
} SingVec4;

typedef struct SingUniforms {
    SingVec4 cam_pos;   /*   0 */
    SingVec4 cam_fwd;   /*  16 */
    SingVec4 cam_right; /*  32 */
    SingVec4 cam_up;    /*  48 */
    float mass_M;       /*  64 */
    float spin_a;       /*  68 */
    float rs;           /*  72 */
    float tan_half_fov; /*  76 */
    float disc_r_inner; /*  80 */
    float disc_r_outer; /*  84 */
    float disc_peak_T;  /*  88 */
    float aspect;       /*  92 */
    float h_step;       /*  96 */
    float escape_r;     /* 100 */
    float horizon_cut;  /* 104 */
    float time_sec;     /* 108 */
    uint32_t width;     /* 112 */
    uint32_t height;    /* 116 */
    uint32_t metric_type; /* 120 */
    uint32_t flags;     /* 124 */
    uint32_t max_steps; /* 128 */
    uint32_t frame_index; /* 132 */
    uint32_t supersample; /* 136: samples per pixel, 1..1024 */
    uint32_t pad_b;     /* 140 */
    float exposure;     /* 144 */
    float bloom_threshold; /* 148 */
    float bloom_strength;  /* 152 */
    float disc_turbulence; /* 156 */
} SingUniforms;

/* uniforms.h:41-52 */
#define SING_FLAG_DOPPLER_ON (1u << 0)
#define SING_FLAG_REDSHIFT_ON (1u << 1)
#define SING_FLAG_DISC_ON (1u << 2)
#define SING_FLAG_STARFIELD_ON (1u << 3)
#define SING_FLAG_ADAPTIVE_STEP (1u << 4)
#define SING_METRIC_SCHWARZSCHILD 0u
#define SING_METRIC_KERR 1u

/* Per-pixel termination class (class of the pixel's first sample). */
#define SING_CLASS_HORIZON 0u  /* captured, including non-finite radius */
#define SING_CLASS_ESCAPE 1u   /* r > escape_r: starfield */
#define SING_CLASS_DISC 2u     /* crossed the equator inside the annulus */
#define SING_CLASS_BUDGET 3u   /* max_steps exhausted */
#define SING_CLASS_INITFAIL 4u /* Kerr null quadratic had no real root at the camera */

/* --- status codes ------------------------------------------------------- */
#define SING_OK 0
#define SING_ERR_ARG 1    /* invalid argument (null pointer, size, range) */
#define SING_ERR_HIP 2    /* HIP runtime error */
#define SING_ERR_STATE 3  /* call not valid in the context's state */
#define SING_ERR_NOMEM 4  /* device allocation failed */
#define SING_ERR_IO 5     /* file could not be read or written */
#define SING_ERR_PARSE 6  /* scene file parse error */
#define SING_ERR_KEY 7    /* scene file unknown key */

/* --- precision of the trace ---------------------------------------------- */
#define SING_PREC_FP32 0u  /* reference arithmetic in FP32 */
#define SING_PREC_FP64 1u  /* the same discrete algorithm in FP64 */
#define SING_PREC_MIXED 2u /* FP32 bulk, FP64 for rays in the photon-sphere band */

/* --- output selection (bit mask) ------------------------------------------ */
#define SING_OUT_RGBA8 (1u << 0)   /* reference semantics: mean of per-sample ACES, sRGB, u8 */
#define SING_OUT_HDR (1u << 1)     /* float RGB: mean of per-sample radiance (before exposure) */
#define SING_OUT_CLASS (1u << 2)   /* u8 termination class of sample 0 */
#define SING_OUT_STEPS (1u << 3)   /* u32 RK4 steps summed over the pixel's samples */
#define SING_OUT_HDR16 (1u << 4)   /* half RGBA HDR (post-pass input) */

/* --- post pass selection --------------------------------------------------- */
#define SING_POST_NONE 0u
#define SING_POST_BLOOM 1u /* half HDR -> extract -> 7-tap H/V -> composite+exposure -> ACES -> sRGB into RGBA8 */

/* --- sub-pixel sampler ----------------------------------------------------- */
#define SING_SAMPLER_HALTON 0u /* supersample Halton(2,3) offsets, mean of per-sample ACES
                                  (render/cuda/kernels/geodesic_kernel.cu:482-502) */
#define SING_SAMPLER_GRID 1u   /* ss x ss grid, ss = clamp(supersample, 1, 4), ACES of the mean radiance
                                  (render/metal/shaders/geodesic_kernel.metal:416-433, blit.metal:54-68) */

/* --- precision refinements of the fixed-step Kerr march (bit mask) ------------
 * Neither has a counterpart in the reference; both change its discrete
 * algorithm, and the oracle implements the same rules. */
#define SING_REFINE_CROSS64 (1u << 0)   /* disc crossing at the root of the FP64 cubic Hermite dense output of
                                           the step (three Newton iterations), not linear in theta
                                           (shared_shader/disc_intersection.h:30-44); phi from the same output */
#define SING_REFINE_REPROJECT (1u << 1) /* after every RK4 step, one Newton step of the null constraint
                                           2H = 0 (shared_shader/kerr_hamilton.h:67-94) along its gradient
                                           in (p_r, p_theta) */

/* Options beyond the 160-byte block.  Zero-initialised = reference semantics,
 * FP32, RGBA8 only. */
typedef struct SingRenderOpts {
    uint32_t precision;   /* SING_PREC_* */
    uint32_t outputs;     /* SING_OUT_* mask; 0 means SING_OUT_RGBA8 */
    uint32_t post;        /* SING_POST_* */
    uint32_t turb_seed;   /* seeded turbulence band phases; 0 = reference bands */
    float adapt_tol;      /* > 0: error-controlled RK4 (step doubling) with this per-step tolerance */
    float band_width;     /* mixed precision: photon-orbit margin below which a ray is traced in FP64 (0 = 0.03) */
    float peak_T_scale;   /* multiplies disc_peak_T; 0 is read as 1 */
    uint32_t step_batch;  /* RK4 steps between lane-refill checks, rounded up to a multiple of 8; the
                             mixed form's FP64 list pass uses a quarter of it (at least 8, a multiple
                             of 8).  0 = tuned defaults: 64 for the FP32 passes and the
                             error-controlled FP64 pass, 32 for the fixed-step FP64 pass; the FP64
                             list 16 (8 with error control) */
    uint32_t waves_per_cu;/* persistent grid size in waves per CU (0 = tuned default) */
    uint32_t sampler;     /* SING_SAMPLER_* */
    uint32_t refine;      /* SING_REFINE_* mask; Kerr with the fixed step (adapt_tol = 0) only */
    uint32_t overlap_chunks; /* 1: a mixed-precision frame of several sample chunks runs each chunk's
                                FP64 list pass and shade on a second stream beside the next chunk's
                                set-up and FP32 pass (same frame bits); 0 (default): chunk after chunk
                                on one stream, which with the FP64 list's longest-first order is as
                                fast or faster.  FP32 and FP64 frames always run serially. */
    uint32_t chunk_samples; /* samples per chunk (0 = as many as fit 2 GiB of 16-byte ray records);
                               smaller chunks hold less device memory and give a mixed frame more
                               chunks to overlap; the frame bits do not depend on it */
    uint32_t reserved[3];
} SingRenderOpts;

typedef struct SingRenderStats {
    double kernel_ms;     /* the frame on the GPU, first kernel start to last kernel end (HIP events) */
    double trace_ms;      /* trace kernels (FP32 pass and FP64 pass), summed over the chunks; with
                             overlap_chunks a chunk's FP64 pass and shade share the time of the
                             next chunk's FP32 pass, so the parts can add up to more than kernel_ms */
    double shade_ms;      /* shade/accumulate kernels */
    double post_ms;       /* post-pass kernels (0 when no post pass) */
    uint64_t rays;        /* pixel samples traced */
    uint64_t steps;       /* RK4 steps executed over all rays */
    uint64_t fp64_rays;   /* rays traced in FP64 */
    uint64_t chunks;      /* sample chunks the frame was split into */
    uint64_t rejected;    /* error-controlled stepping: rejected step attempts (each is 3 RK4 steps of work;
                             an accepted step is also 3) */
    double trace32_ms;    /* the FP32 trace kernels alone (HIP events around each launch) */
    double trace64_ms;    /* the FP64 trace kernels alone */
    uint64_t steps32;     /* RK4 steps executed by the FP32 trace kernels, including rays later handed
                             to the FP64 pass */
    uint64_t steps64;     /* RK4 steps executed by the FP64 trace kernels */
} SingRenderStats;

typedef struct SingContext SingContext;

/* --- lifecycle (RenderBackend::initialize/shutdown/resize) ---------------- */
/* Creates a context for width x height on HIP device `device`; NULL on failure
 * (reason from sing_last_error(NULL)). */
SingContext* sing_create(uint32_t width, uint32_t height, int device);
void sing_destroy(SingContext* ctx);
int sing_resize(SingContext* ctx, uint32_t width, uint32_t height);
/* Last error message for ctx, or for ctx == NULL the calling thread's last
 * error (each host thread has its own, like errno).  Contexts are independent:
 * several host threads may each drive their own context at the same time; one
 * context is used by one thread at a time. */
const char* sing_last_error(const SingContext* ctx);
const char* sing_backend_name(void); /* "HIP-gfx950" (RenderBackend::name) */
int sing_abi_version(void);

/* --- render (RenderBackend::render_frame + the kernel launch) --------------
 * Replaces CudaBackend::render_frame (render/cuda/cuda_backend.cu:139-152: one
 * kernel launch, then cudaDeviceSynchronize).  Synchronous: returns after the
 * frame is complete.  `u` is copied at call time.  u->width/height must equal
 * the context size.  Equivalent to sing_render_async followed by sing_wait. */
int sing_render(SingContext* ctx, const SingUniforms* u, const SingRenderOpts* opts, SingRenderStats* stats);
/* The same frame without the wait: validates the arguments (errors are
 * returned here, as from sing_render), enqueues the frame on the context's
 * streams and returns; `u` and `opts` may be freed on return.  The host is
 * free until sing_wait.  Any other call on the context (capture, render,
 * resize, post pass, batch, differentiable render, diagnostics, destroy)
 * first completes a pending frame, so its outputs are never read half-done;
 * the frame's stats stay available to sing_wait until the next submission.
 * A capture or post pass that completes a frame which failed returns that
 * frame's error (sing_wait returns it too). */
int sing_render_async(SingContext* ctx, const SingUniforms* u, const SingRenderOpts* opts);
/* Waits for the frame submitted by sing_render_async (if it has not completed
 * yet) and returns its status and stats (optional), as sing_render would
 * have.  SING_ERR_STATE when no frame was submitted since the last wait. */
int sing_wait(SingContext* ctx, SingRenderStats* stats);

/* --- capture (RenderBackend::capture_frame) ---------------------------------
 * Copies the last frame's outputs into caller buffers of at least the stated
 * byte size: RGBA8 W*H*4, HDR W*H*3*4, class W*H, steps W*H*4, HDR16 W*H*8. */
int sing_capture_rgba8(SingContext* ctx, uint8_t* dst, size_t bytes);
int sing_capture_hdr(SingContext* ctx, float* dst, size_t bytes);
int sing_capture_class(SingContext* ctx, uint8_t* dst, size_t bytes);
int sing_capture_steps(SingContext* ctx, uint32_t* dst, size_t bytes);
int sing_capture_hdr16(SingContext* ctx, uint16_t* dst, size_t bytes);

/* Re-runs the bloom post pass (extract, 7-tap blurs, composite with exposure,
 * ACES, sRGB; render/vulkan/shaders/bloom.hlsl:8-77, blit.hlsl:43-58) over the
 * last frame's half-float HDR (a frame rendered with SING_OUT_HDR16 or bloom)
 * with the exposure, bloom_threshold and bloom_strength of *u (nothing else of
 * *u is read), into the RGBA8 image: a new exposure or bloom setting without
 * re-tracing.  The pass is launched `repeat` times back to back (1..1000;
 * every launch writes the same image); stats (optional) gets post_ms = the
 * mean time per launch between HIP events around all of them and kernel_ms =
 * their total, other fields zero.  No counterpart in the reference, whose post
 * passes run inside each frame. */
int sing_post_bloom(SingContext* ctx, const SingUniforms* u, uint32_t repeat, SingRenderStats* stats);

/* --- render without device allocation ---------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the sing_render pipeline with the same kernels, options and
 * results, writing straight into caller-owned device buffers, with every
 * working buffer taken from one static 256 MiB device array per GPU (held until
 * the process ends).  No context; no hipMalloc/hipFree.  Frames whose working
 * set for one sample per chunk exceeds the array (about 3.7 Mpixel, e.g.
 * 2560x1440, without bloom) are rejected with SING_ERR_ARG; sing_render has no
 * such limit.  Synchronous on `stream` (NULL = the null stream) of `device`,
 * which becomes the calling thread's current device.  The first call on a
 * device also looks up the array and uploads the 8 KiB sample-offset tables
 * (make it before any graph capture).  Calls that overlap in time on one GPU
 * share the array and race.  Output pointers: rgba W*H*4, hdr W*H*3 floats,
 * hdr16 W*H*4 halves (16-byte aligned), cls W*H, steps W*H u32, each required
 * when its SING_OUT_* bit (or bloom, for rgba) is selected.  stats: rays, steps,
 * fp64_rays, chunks, rejected (timings are 0).  Errors: sing_last_error(NULL). */
typedef struct SingDeviceOutputs {
    uint8_t* rgba;
    float* hdr;
    uint16_t* hdr16;
    uint8_t* cls;
    uint32_t* steps;
} SingDeviceOutputs;
int sing_render_noalloc(int device, const SingUniforms* u, const SingRenderOpts* opts, const SingDeviceOutputs* out,
                        void* stream, SingRenderStats* stats);

/* --- host-side scene ingestion (no GPU needed) ------------------------------ */
/* render_backend.hpp:46-82 Scene (the fields cuda_cli uses) */
typedef struct SingScene {
    uint32_t metric;          /* SING_METRIC_* */
    float mass;               /* mass_solar */
    float spin_a_over_M;
    float disc_inner_M;
    float disc_outer_M;
    uint32_t doppler_on;
    uint32_t redshift_on;
    uint32_t disc_on;
    uint32_t supersample;     /* samples per pixel */
    float exposure;
    float bloom_threshold;
    float bloom_strength;
    float disc_turbulence;
    float disc_peak_T_K;
    uint32_t max_steps;       /* 0 = backend default (2500) */
    float h_step;             /* 0 = backend default (0.4) */
    float escape_r_M;         /* 0 = backend default (200 M) */
    float horizon_cut_rs;     /* 0 = backend default (1.02 rs) */
} SingScene;

/* render_backend.hpp:84-99 CameraState */
typedef struct SingCamera {
    float position[3];
    float basis[9];           /* rows: right, up, -forward */
    float fov_y_radians;
} SingCamera;

/* scene_config.hpp:39-58 SceneConfig */
typedef struct SingSceneConfig {
    float mass;
    float spin;
    float disc_r_inner;       /* -1 = ISCO (the loader keeps the raw value) */
    float disc_r_outer;
    uint32_t disc_doppler_on;
    uint32_t disc_redshift_on;
    float camera_fov_deg;
    float camera_distance;
    float camera_elevation;
    float h_step;
    int32_t max_steps;
} SingSceneConfig;

/* Scene{} defaults (render_backend.hpp:52-74). */
void sing_scene_defaults(SingScene* out);
/* SceneConfig{} defaults (scene_config.hpp:39-58). */
void sing_scene_config_defaults(SingSceneConfig* out);
/* Parses `key = value` text (scene_config.hpp:134-191).  On failure returns
 * SING_ERR_PARSE or SING_ERR_KEY, with *line (1-based) and the offending
 * line/key copied into errbuf. */
int sing_parse_scene_config(const char* text, SingSceneConfig* out, int* line, char* errbuf, size_t errbuf_len);
/* scene_config.hpp:193-199; SING_ERR_IO when the file cannot be opened. */
int sing_load_scene_config(const char* path, SingSceneConfig* out, int* line, char* errbuf, size_t errbuf_len);
/* cuda_cli/main.cpp:193-217: SceneConfig -> Scene (metric = Kerr iff spin > 0). */
void sing_scene_from_config(const SingSceneConfig* cfg, SingScene* scene);
/* cuda_cli/main.cpp:68-109 */
void sing_orbital_camera(float distance_M, float elevation_rad, float azimuth_rad, float fov_y_deg, SingCamera* out);
/* cuda_backend.cu:45-101 */
void sing_pack_uniforms(const SingScene* scene, const SingCamera* cam, uint32_t width, uint32_t height,
                        float time_sec, SingUniforms* out);
/* Kerr closed forms (core/include/physics/kerr.hpp:36-108), double precision:
 * which = 0 r+, 1 r-, 2 ISCO prograde, 3 ISCO retrograde, 4 photon orbit prograde,
 * 5 photon orbit retrograde. */
double sing_kerr_radius(int which, double mass, double spin_a);

/* --- differentiable render (discrete adjoint) ------------------------------
 * No counterpart in the reference (nothing there is differentiable).  The scene
 * is Kerr with parameters p = (a/M, inclination, r_in/M, r_out/M, peak-T
 * scale); the camera is the orbital camera at elevation pi/2 - inclination;
 * the disc has soft annulus edges of width soft_edge (M) so the image is
 * differentiable in r_in and r_out.  One sample per pixel at the pixel centre.
 * `base` supplies resolution (must equal the context), mass, h_step,
 * max_steps, escape_r, flags, peak T, turbulence and time; its camera and
 * spin/disc fields are ignored. */
typedef struct SingDiffParams {
    double spin;      /* a/M */
    double incl;      /* inclination (rad): camera polar angle */
    double r_in;      /* disc inner radius / M */
    double r_out;     /* disc outer radius / M */
    double t_scale;   /* multiplies base->disc_peak_T */
} SingDiffParams;

typedef struct SingDiffCamera {
    double distance;  /* M */
    double azimuth;   /* rad */
    double fov_y_deg;
} SingDiffCamera;

typedef struct SingDiffOpts {
    uint32_t precision;        /* SING_PREC_*: FP32, FP64, or mixed (FP64 for band rays) */
    uint32_t checkpoint_every; /* RK4 steps between stored states (0 = 64) */
    double soft_edge;          /* annulus edge ramp width in M (0 = 1.0) */
    uint32_t r_in_isco;        /* 1: r_in = prograde ISCO of a/M (p->r_in ignored); grad[0] is then the
                                  total derivative through r_in(a), grad[2] the partial dL/d(r_in/M) */
    uint32_t target_per_eval;  /* sing_diff_loss_grad_batch: 1 = target (and weights) hold one image per
                                  parameter point, n of them one after another (with one camera per point:
                                  several views of one scene); 0 = one image shared by all points */
    uint32_t reserved[4];
} SingDiffOpts;

typedef struct SingDiffStats {
    double forward_ms;
    double backward_ms;        /* adjoint sweep + reduction */
    uint64_t rays;
    uint64_t steps;            /* forward RK4 steps */
    uint64_t fp64_rays;
    uint64_t reserved[3];
} SingDiffStats;

/* Forward only: radiance (W*H*3 float, host) of the differentiable model. */
int sing_diff_forward(SingContext* ctx, const SingUniforms* base, const SingDiffParams* p, const SingDiffCamera* cam,
                      const SingDiffOpts* opts, float* hdr_out, SingDiffStats* stats);
/* Loss = sum_pixels w_p |I_p - target_p|^2 (target W*H*3 float, weights W*H
 * float or NULL for 1) and its gradient wrt the 5 parameters (grad[5]);
 * hdr_out (optional) receives the radiance. */
int sing_diff_loss_grad(SingContext* ctx, const SingUniforms* base, const SingDiffParams* p,
                        const SingDiffCamera* cam, const SingDiffOpts* opts, const float* target,
                        const float* weights, double* loss, double* grad, float* hdr_out, SingDiffStats* stats);

/* n parameter points (1..SING_DIFF_MAX_BATCH) of the same image in one set of
 * launches: the losses (loss[n]) and gradients (grad[n*5]) that n calls of
 * sing_diff_loss_grad would return, bit for bit.  cams: one camera for all
 * (n_cams = 1) or one per point (n_cams = n); target W*H*3 floats (weights
 * W*H or NULL) shared by all points, or with opts->target_per_eval = 1 one
 * image per point, n of them one after another (several views of one scene:
 * one camera and one observed image per point, the fit's loss their sum);
 * hdr_out (optional) n*W*H*3 floats, image after image.  One 512x512
 * evaluation leaves most lanes idle
 * (its time is its longest ray's), so independent evaluations — finite-
 * difference stencils, multi-start or population fits, several cameras —
 * share the launches.  No counterpart in the reference. */
#define SING_DIFF_MAX_BATCH 32u
int sing_diff_loss_grad_batch(SingContext* ctx, const SingUniforms* base, uint32_t n, const SingDiffParams* params,
                              const SingDiffCamera* cams, uint32_t n_cams, const SingDiffOpts* opts,
                              const float* target, const float* weights, double* loss, double* grad, float* hdr_out,
                              SingDiffStats* stats);

/* The backward pass for a caller's own loss: given dl_dhdr = dLoss/dI (W*H*3
 * floats, one RGB cotangent per pixel of the HDR image I), returns grad[5] =
 * sum_p dl_dhdr_p . dI_p/dparams (the vector-Jacobian product of the render,
 * same parameters and model as sing_diff_loss_grad) and *pairing (optional) =
 * sum_p dl_dhdr_p . I_p (I before its rounding to the float image), the
 * scalar whose gradient that is.  With dl_dhdr =
 * 2 w (I - target) it is sing_diff_loss_grad's gradient: the backward pass
 * takes a target image there and a dLoss/dHDR image here.  No counterpart in
 * the reference (it has no differentiable path). */
int sing_diff_vjp(SingContext* ctx, const SingUniforms* base, const SingDiffParams* p, const SingDiffCamera* cam,
                  const SingDiffOpts* opts, const float* dl_dhdr, double* pairing, double* grad, float* hdr_out,
                  SingDiffStats* stats);

/* Per-pixel results of the last sing_diff_* call on ctx (either may be NULL);
 * after a batch call, those of its first parameter point:
 * grad_out W*H*5 doubles, the pixel's loss gradient wrt the 5 parameters (r_in
 * partial; zero after sing_diff_forward); meta_out W*H*4 u32: RK4 steps,
 * accepted disc crossings, termination (1 horizon, 2 escape, 3 disc, 4 budget,
 * 5 init fail) and the precision the pixel ran in (0 FP32, 1 FP64). */
int sing_diff_capture_pixels(SingContext* ctx, double* grad_out, uint32_t* meta_out);

/* --- batched render ------------------------------------------------------------
 * No counterpart in the reference (cuda_cli renders one scene per launch,
 * cuda_cli/main.cpp:243-299).  Renders n scenes of any sizes and sample counts
 * (each a 160-byte SingUniforms, semantics as sing_render) with one persistent
 * trace launch over all of them: each wavefront works on one scene at a time
 * and moves to another scene with work left when its own runs dry.  Outputs
 * (reference RGBA8 semantics, optional HDR and class) go to caller host buffers
 * per scene.  The context's own frame size is unrelated to the batch's sizes. */
typedef struct SingBatchItem {
    const SingUniforms* uniforms;
    uint8_t* rgba;   /* W*H*4, or NULL */
    float* hdr;      /* W*H*3 mean radiance before exposure, or NULL */
    uint8_t* cls;    /* W*H class of sample 0, or NULL */
} SingBatchItem;

typedef struct SingBatchOpts {
    uint32_t precision;     /* SING_PREC_* */
    uint32_t use_graph;     /* 1: record the batch's launches once as a hipGraph and replay it for batches
                               of the same layout (sizes, sample counts, metrics, options) */
    uint32_t autotune;      /* 1: time candidates for waves per CU, step batch, rays per refill and queue
                               chunk on a 1/8 sample of this batch and keep the fastest for this (metric,
                               precision, stepping) in the context */
    uint32_t step_batch;    /* RK4 steps between lane refills (0 = tuned or 64; rounded up to a
                               multiple of 8) */
    uint32_t waves_per_cu;  /* persistent waves per CU (0 = tuned or 24) */
    float adapt_tol;        /* > 0: error-controlled stepping (applies to the Kerr scenes) */
    uint32_t turb_seed;     /* seeded disc-band phases, as SingRenderOpts */
    uint32_t record_budget_mib; /* device memory for one group's termination records (0 = 32 GiB); scenes are
                                   split into groups that fit, traced one group after another */
    uint32_t rays_per_refill;   /* a wave claims new rays once at least this many of its lanes are idle (or
                                   none is live); 0 = tuned or 32, at most 64 */
    uint32_t queue_chunk;       /* work items a wave takes from its scene's queue per atomic and hands to its
                                   lanes as they free up; 1 = one per idle lane; 0 = tuned, else 256
                                   (64 below 1024 rays per wave of the grid, 1 below 256) */
    uint32_t reserved[2];
} SingBatchOpts;

typedef struct SingBatchStats {
    double kernel_ms;       /* all kernels of the batch (one graph launch when use_graph) */
    double copy_ms;         /* device -> host copies of the outputs */
    uint64_t rays;
    uint64_t steps;
    uint64_t fp64_rays;
    uint64_t rejected;
    uint32_t step_batch;    /* launch shape used */
    uint32_t waves_per_cu;
    uint32_t graph_replayed;   /* 1 if an existing graph was replayed */
    uint32_t groups;           /* record-memory groups the batch was split into */
    uint32_t rays_per_refill;  /* claim policy used */
    uint32_t queue_chunk;
} SingBatchStats;

int sing_render_batch(SingContext* ctx, const SingBatchItem* items, uint32_t n, const SingBatchOpts* opts,
                      SingBatchStats* stats);

/* --- the CPU driver's diagnostics (cli/main.cpp:1269-1621) -------------------
 * kerr-geometry, photon-orbit and benchmark of singularity_cli, in float like
 * the reference.  The geometry is closed forms on the host (no device needed);
 * the orbit and the ray fan are set up on the host exactly as the driver does
 * and integrated on the device by the renderer's RK4 steppers. */

/* cli/main.cpp:1415-1476 (core/include/physics/kerr.hpp:36-108). */
typedef struct SingKerrGeometry {
    float mass, spin, spin_over_mass;
    float outer_horizon, inner_horizon;
    float ergosphere_polar, ergosphere_equatorial;
    float photon_sphere_prograde, photon_sphere_retrograde;
    float isco_prograde, isco_retrograde;
} SingKerrGeometry;
/* SING_ERR_ARG (message in sing_last_error(NULL)) for mass <= 0 or spin
 * outside [0, mass]. */
int sing_kerr_geometry(float mass, float spin, SingKerrGeometry* out);

/* cli/main.cpp:1279-1410: one equatorial circular null orbit at r_init for
 * `orbits` periods with RK4 step h_step (Schwarzschild when spin = 0, else the
 * prograde Kerr orbit with E = 1, L_z = (r sqrt(Delta) - 2 M a)/(r - 2M)).
 * Rows are those of the driver's CSV: the start, every
 * max(1, int(period / (200 h))) steps and the last step, or a row with
 * step = -1 where the ray leaves (non-finite r or r < 1.001 r+).  When cap is
 * below the rows the orbit can produce, returns SING_ERR_ARG with that bound in
 * *n_rows.  Argument errors (the driver's exit-2 cases) give SING_ERR_ARG. */
typedef struct SingOrbitRow {
    int32_t step;
    float t, r, theta, phi;
} SingOrbitRow;
int sing_photon_orbit(SingContext* ctx, float mass, float spin, float r_init, float orbits, float h_step,
                      SingOrbitRow* rows, uint32_t cap, uint32_t* n_rows, double* kernel_ms);

/* cli/main.cpp:1492-1590: the benchmark's equatorial ray fan (M = 1,
 * h = 0.1): ray i has impact parameter b_min + (b_max - b_min)(i + 1/2)/n_rays
 * and steps until max_steps, a non-finite r or phi, or r below the cut
 * (1.01 r_s Schwarzschild, 1.02 r+ Kerr).  One device lane per ray. */
typedef struct SingFanOpts {
    uint32_t metric;      /* SING_METRIC_SCHWARZSCHILD or SING_METRIC_KERR */
    float spin;           /* a/M, Kerr only: [0, 1) */
    uint32_t n_rays;      /* >= 1 */
    uint32_t max_steps;
    float b_min, b_max;   /* b_min < b_max, in M */
    uint32_t reserved[2];
} SingFanOpts;
typedef struct SingFanStats {
    uint64_t total_steps; /* RK4 steps over all rays */
    double kernel_ms;     /* the integration kernel (HIP events) */
    double setup_ms;      /* host set-up and upload of the initial states */
    uint64_t reserved;
} SingFanStats;
/* steps_per_ray: n_rays counts, or NULL. */
int sing_benchmark_fan(SingContext* ctx, const SingFanOpts* opts, uint32_t* steps_per_ray, SingFanStats* stats);

#ifdef __cplusplus
}
#endif

#endif /* SING_ABI_H */
