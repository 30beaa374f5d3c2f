/*
 * seis3d.h - C ABI of the MI355X-native 3D elastic velocity-stress solver.
 *
 * Every entry point replaces one function of the reference's Seis3D solver
 * API (cuda/Cuda_Seis3D.cuh:6-42 on the POD types of include/Common3D.h:24-98)
 * or one step of its session driver (src/Seis3D_Sim.cpp,
 * src/Seis3D/Acquisition3D.cpp).  The reference API is a set of C++ free
 * functions that throw; here every function returns an error code and keeps a
 * thread-local message (seis3d_last_error).  include/Seis3D.hpp maps the codes
 * back to the reference's exception types and message prefixes.
 *
 * Error codes: 0 ok, 1 invalid argument (std::invalid_argument), 2 HIP/runtime
 * failure (std::runtime_error), 3 lifecycle misuse (std::logic_error),
 * 4 non-finite wavefield detected (std::runtime_error).
 *
 * Layouts at the boundary are the reference's: padded grid gx*gy*gz with x
 * fastest (include/Common3D.h:96-98); material planes are interleaved SI
 * triples [Vp,Vs,rho] with x fastest (cuda/Cuda_Seis3D.cu:696-711); receiver
 * samples are [receiver][vx,vy,vz] (cuda/Cuda_Seis3D.cu:421-428) and gathers
 * [sample][receiver][vx,vy,vz] (include/Seis3D_Sim.h:52).
 */
#ifndef SEIS3D_H
#define SEIS3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum seis3d_status {
    SEIS3D_OK = 0,
    SEIS3D_EINVAL = 1,
    SEIS3D_EHIP = 2,
    SEIS3D_ELIFECYCLE = 3,
    SEIS3D_ENONFINITE = 4
};

/* Field3D (include/Common3D.h:18). */
enum seis3d_field {
    SEIS3D_VX = 0, SEIS3D_VY, SEIS3D_VZ,
    SEIS3D_SXX, SEIS3D_SYY, SEIS3D_SZZ, SEIS3D_SXY, SEIS3D_SYZ, SEIS3D_SXZ,
    /* Material arrays, readable through seis3d_download_field. */
    SEIS3D_INV_RHO = 9, SEIS3D_LAMBDA = 10, SEIS3D_MU = 11
};

/* ModelKind3D (include/Common3D.h:20-21), same numeric IDs. */
enum seis3d_model_kind {
    SEIS3D_MODEL_UNIFORM = 0, SEIS3D_MODEL_TWO_LAYER, SEIS3D_MODEL_SPHERE, SEIS3D_MODEL_IMPORTED,
    SEIS3D_MODEL_THREE_LAYER, SEIS3D_MODEL_DIPPING_LAYER, SEIS3D_MODEL_LOW_VELOCITY_LAYER,
    SEIS3D_MODEL_FAULT, SEIS3D_MODEL_GRADIENT, SEIS3D_MODEL_SALT_DOME
};

/* SourceKind3D (include/Common3D.h:22). */
enum seis3d_source_kind { SEIS3D_EXPLOSION = 0, SEIS3D_FORCE = 1 };

/* SimulationContext3D (include/Common3D.h:24-42); bools are int32 0/1. */
typedef struct seis3d_ctx {
    int32_t nx, ny, nz;
    int32_t pml;
    int32_t free_surface;
    int32_t reflecting;
    float spacing, cfl, dt;
    int32_t manual_dt;
    float requested_dt_ms;
    float safe_dt;
    float vp, vs, density;
    float layer_vp, layer_vs, layer_density;
    float max_vp, frequency;
    int32_t model_kind;
    float interface_depth, layer_dip, layer_thickness;
    float fault_throw, anomaly_radius;
    /* Derived by seis3d_validate_ctx only. */
    int32_t gx, gy, gz, ox, oy, oz;
    uint64_t cells;
} seis3d_ctx;

/* Source3D (include/Common3D.h:44-49). */
typedef struct seis3d_source {
    int32_t x, y, z;
    int32_t kind;
    float amplitude;
    float direction[3];
} seis3d_source;

/* Receiver3D (include/Common3D.h:51). */
typedef struct seis3d_receiver { int32_t x, y, z; } seis3d_receiver;

/* MemoryBudget3D (include/Common3D.h:72-74). */
typedef struct seis3d_budget {
    uint64_t field_bytes, cpml_bytes, view_bytes, total_bytes;
} seis3d_budget;

/* Survey3D receiver/shot spread (include/Seis3D_Sim.h:36-45). */
typedef struct seis3d_survey {
    int32_t receiver_count_x, receiver_count_y;
    int32_t receiver_origin[3];
    int32_t receiver_spacing[2];
    int32_t shot_count_x, shot_count_y;
    int32_t shot_origin[3];
    int32_t shot_spacing[2];
} seis3d_survey;

/* Per-kernel timing of a run (HIP events on the session stream). */
typedef struct seis3d_timing {
    double velocity_ms;  /* summed duration of the fused velocity half-step kernel */
    double stress_ms;    /* summed duration of the fused stress half-step kernel */
    double total_ms;     /* whole run, first launch to last completion */
    int64_t launches;    /* launches of each of the two kernels */
} seis3d_timing;

typedef struct seis3d_session seis3d_t;

/* ---- Context, validation, budgets (host only, no GPU needed) ---------- */

/* Defaults of SimulationContext3D{} (include/Common3D.h:24-42). */
void seis3d_default_ctx(seis3d_ctx* ctx);
/* Defaults of Source3D{} (include/Common3D.h:44-49). */
void seis3d_default_source(seis3d_source* src);
/* Defaults of Survey3D{} (include/Seis3D_Sim.h:36-45). */
void seis3d_default_survey(seis3d_survey* survey);
/* validateContext3D (cuda/Cuda_Seis3D.cu:433-483). */
int seis3d_validate_ctx(seis3d_ctx* ctx);
/* estimateMemory3D (cuda/Cuda_Seis3D.cu:504-516): the reference's formula. */
int seis3d_estimate_memory(const seis3d_ctx* ctx, seis3d_budget* out);
/* Device bytes this library allocates for ctx (fields, materials, CPML). */
int seis3d_device_bytes(const seis3d_ctx* ctx, uint64_t* out);
/* makeCoefficients3D (cuda/Cuda_Seis3D.cu:45-64): [half][n][a,b,1/kappa]. */
int seis3d_cpml_coefficients(const seis3d_ctx* ctx, int axis, float* out, size_t capacity);
/* fitSurvey3D (src/Seis3D_Sim.cpp:42-60). */
int seis3d_fit_survey(const seis3d_ctx* ctx, seis3d_survey* survey);
/* configureReceivers3D layout (src/Seis3D/Acquisition3D.cpp:21-37); *count in/out. */
int seis3d_survey_receivers(const seis3d_ctx* ctx, const seis3d_survey* survey,
                            seis3d_receiver* out, int* count);
/* positionSequenceShot3D (src/Seis3D/Acquisition3D.cpp:233-242). */
int seis3d_sequence_shot(const seis3d_ctx* ctx, const seis3d_survey* survey, int index,
                         seis3d_source* src);
/* Gather timing of startGather3D (src/Seis3D/Acquisition3D.cpp:201-209). */
int seis3d_gather_timing(const seis3d_ctx* ctx, int stride, int* interval_us,
                         double* first_sample_time);
/* Ricker sample of sourceKernel3D (cuda/Cuda_Seis3D.cu:319-320) at time t. */
float seis3d_ricker(float frequency, double time);

/* ---- Session lifetime (initializeGpu3D / releaseGpu3D) ------------------ */

/* initializeGpu3D (cuda/Cuda_Seis3D.cu:519-555) on the current HIP device.
 * ctx must have passed seis3d_validate_ctx. */
int seis3d_create(const seis3d_ctx* ctx, seis3d_t** out);
/* releaseGpu3D (cuda/Cuda_Seis3D.cu:558-569); NULL is accepted. */
void seis3d_destroy(seis3d_t* t);
/* The context the session was created with (after any refresh). */
int seis3d_get_ctx(const seis3d_t* t, seis3d_ctx* out);
/* resetGpu3D (cuda/Cuda_Seis3D.cu:572-577). */
int seis3d_reset(seis3d_t* t);
/* refreshCpmlCoefficients3D (cuda/Cuda_Seis3D.cu:580-599). */
int seis3d_refresh_cpml(seis3d_t* t, const seis3d_ctx* ctx);
/* Read back the device CPML table of one axis ([half][n][a,b,1/kappa]). */
int seis3d_read_cpml(seis3d_t* t, int axis, float* out, size_t capacity);

/* Learned CPML profile.  No counterpart in the reference, whose profile is fixed
 * (makeCoefficients3D, cuda/Cuda_Seis3D.cu:45-64: sigma = sigma_max d^3 with
 * sigma_max = -4 vp_max ln(1e-6) / (2 pml h), kappa = 1 + 2 d^3, alpha = pi f (1 - d)
 * at layer depth d in [0, 1]).  The learned profile keeps the table layout and the
 * a/b/kappa formulas and replaces the three shapes by
 *     sigma = e^p[0] sigma_max d^p[1],  kappa = 1 + (p[2] - 1) d^p[1],
 *     alpha = p[3] pi f (1 - d)^p[4]   (alpha = 0 at d = 0, as the reference),
 * so the step costs exactly what the reference profile costs.  The reference
 * profile is p = {0, 3, 3, 1, 1}.  Accepted ranges: |p[0]| <= 5, 1 <= p[1] <= 8,
 * 1 <= p[2] <= 20, 0 <= p[3] <= 10, 0.1 <= p[4] <= 8; anything else is
 * SEIS3D_EINVAL.  seis3d_set_cpml_profile uploads the tables for the session (NULL
 * restores the reference profile); seis3d_refresh_cpml keeps the session's profile.
 * Refused (SEIS3D_ELIFECYCLE) while a batch holds the session or a learned frame
 * is attached; a no-op in reflecting mode. */
#define SEIS3D_PROFILE_PARAMS 5
int seis3d_set_cpml_profile(seis3d_t* t, const double* profile);
int seis3d_get_cpml_profile(const seis3d_t* t, double* profile, int* learned);
int seis3d_cpml_profile_coefficients(const seis3d_ctx* ctx, const double* profile, int axis, float* out,
                                     size_t capacity);

/* ---- Model ---------------------------------------------------------------- */

/* uploadModelPlane3D (cuda/Cuda_Seis3D.cu:696-711). */
int seis3d_upload_plane(seis3d_t* t, int z, const float* vp_vs_rho);
/* finishModelUpload3D (cuda/Cuda_Seis3D.cu:714-717). */
int seis3d_finish_upload(seis3d_t* t);
/* downloadModelPlane3D (cuda/Cuda_Seis3D.cu:720-733). */
int seis3d_download_plane(seis3d_t* t, int z, float* vp_vs_rho);
/* probeMaterial3D (cuda/Cuda_Seis3D.cu:684-693); out of range leaves values. */
int seis3d_probe(seis3d_t* t, const seis3d_receiver* p, float values[3]);

/* ---- Stepping ------------------------------------------------------------- */

/* runStep3D (cuda/Cuda_Seis3D.cu:602-624): one leapfrog step, enqueue only. */
int seis3d_step(seis3d_t* t, const seis3d_source* src, uint64_t step);
/* n_steps runStep3D calls from first_step, enqueue only.  When a gather is
 * recording, samples are taken after every step s with (s+1) % stride == 0
 * into the device gather, with no host synchronisation
 * (advanceSimulation3D, src/Seis3D_Sim.cpp:120-153). */
int seis3d_run(seis3d_t* t, const seis3d_source* src, uint64_t first_step, int64_t n_steps);
/* Block until the session stream is idle; reports the non-finite flag. */
int seis3d_synchronize(seis3d_t* t);

/* ---- Receivers and gathers ------------------------------------------------ */

/* setReceivers3D (cuda/Cuda_Seis3D.cu:649-668): validated atomic replace. */
int seis3d_set_receivers(seis3d_t* t, const seis3d_receiver* receivers, int count);
int seis3d_receiver_count(const seis3d_t* t);
/* sampleReceivers3D (cuda/Cuda_Seis3D.cu:671-681): [r][vx,vy,vz] in owned
 * pinned memory, valid until the next sample, set_receivers or destroy;
 * *out is NULL when no receivers are set. */
int seis3d_sample_receivers(seis3d_t* t, const float** out);
/* startGather3D recording state (src/Seis3D/Acquisition3D.cpp:188-230):
 * allocate a device gather of target_samples x receivers x 3 and sample
 * every stride steps during seis3d_run. */
int seis3d_gather_begin(seis3d_t* t, int target_samples, int stride);
/* Samples recorded so far (host counter, no synchronisation). */
int seis3d_gather_count(const seis3d_t* t);
/* Copy the recorded [sample][receiver][3] gather to host; checks the
 * non-finite flag first (error 4, as sampleReceivers3D). */
int seis3d_gather_read(seis3d_t* t, float* out, size_t capacity_floats, int* samples);
/* Stop recording and release the device gather. */
int seis3d_gather_end(seis3d_t* t);

/* ---- Wavefield access (snapshots; replaces the GL interop views) --------- */

/* Copy one padded field (seis3d_field) to host in the reference layout
 * gx*gy*gz, x fastest.  Synchronises the session stream. */
int seis3d_download_field(seis3d_t* t, int field, float* out);
/* Overwrite one padded wavefield (0..8) from a host array in that layout. */
int seis3d_upload_field(seis3d_t* t, int field, const float* in);
/* Physical-domain XZ (axis 1), YZ (axis 0) or XY (axis 2) slice at physical
 * index `at`, collocated component as sampleScalar3D (cuda/Cuda_Seis3D.cu:335-352):
 * 0 vz, 1 vx, 2 vy, 3 |v|, 4 pressure, 5 vp, 6 vs, 7 rho.  The slice covers
 * the display domain of makeViewLayout3D (CPML included, ghosts excluded). */
int seis3d_extract_slice(seis3d_t* t, int axis, int at, int component, float* out, size_t capacity);

/* ViewLayout3D (include/Common3D.h:53-60): sample counts of the packed view. */
typedef struct seis3d_view_layout {
    uint64_t slice_offset[3];
    int32_t slice_width[3], slice_height[3];
    uint64_t volume_offset, total_samples;
    int32_t volume_x, volume_y, volume_z;
    int32_t domain_size[3], domain_origin[3];
} seis3d_view_layout;
/* makeViewLayout3D (cuda/Cuda_Seis3D.cu:486-501). */
int seis3d_make_view_layout(const seis3d_ctx* ctx, seis3d_view_layout* out);

/* Statistics of extractViews3D: maximum |wave| (physical cells unless
 * include_boundary), material range over physical cells, maximum |v|. */
typedef struct seis3d_view_stats {
    float maximum, material_min, material_max, velocity_maximum;
} seis3d_view_stats;
/* extractViews3D (cuda/Cuda_Seis3D.cu:627-646): the XZ, YZ and XY planes at
 * slices[0..2] (physical x, y, z) and, if volume, the trilinear volume, as
 * [sample][wave, material] floats in host memory (capacity in floats); wave
 * component 0..7 as seis3d_extract_slice, material component 5..7.  A
 * non-finite sample fails with code 4. */
int seis3d_extract_views(seis3d_t* t, float* out, size_t capacity, const int32_t slices[3], int component,
                         int material_component, int volume, int include_boundary, seis3d_view_stats* stats);
/* No reference counterpart: seis3d_extract_views for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, results and errors, but it calls no hipMalloc/hipFree: the device staging is
 * a fixed array of the library, one per GPU (16 MiB + 256 B, held until the process ends), so
 * at most 2,097,152 samples per call (the three planes alone always fit; a volume may not);
 * a larger request fails with code 1.  Calls that overlap on one GPU share that array and race.
 * The session's device must be current. */
int seis3d_extract_views_noalloc(seis3d_t* t, float* out, size_t capacity, const int32_t slices[3], int component,
                                 int material_component, int volume, int include_boundary,
                                 seis3d_view_stats* stats);

/* ---- Tuning and measurement ---------------------------------------------- */

/* Kernel family: 0 = direct (one thread per cell), 4 = 2.5D LDS-DMA streaming
 * (the default).  tile_y rows (4 or 8) and z_chunk planes (1..128) per
 * workgroup of family 4.  z_chunk 0 with tile_y 8: one round of resident
 * workgroups, each tile column with its own chunk count so that columns in CPML
 * slabs (more work per plane) get shorter chunks; where that does not fit one
 * round, and for out-of-place and batched steps, 16-plane chunks. */
int seis3d_set_tiles(seis3d_t* t, int family, int tile_y, int z_chunk);
int seis3d_get_tiles(const seis3d_t* t, int* family, int* tile_y, int* z_chunk);
/* seis3d_run with HIP events around every fused-kernel launch. */
int seis3d_run_timed(seis3d_t* t, const seis3d_source* src, uint64_t first_step,
                     int64_t n_steps, seis3d_timing* timing);
/* Algorithmic HBM bytes of one step: 4*(30*N_upd + 12*N_psi), with the
 * velocity and stress kernel shares. */
int seis3d_step_bytes(const seis3d_ctx* ctx, uint64_t* total, uint64_t* velocity,
                      uint64_t* stress, uint64_t* n_upd, uint64_t* n_psi);

/* ---- Autotuner and multi-shot batches (no reference counterpart) ---------- */

/* Time each candidate kernel shape for `steps` steps on a scratch session of
 * the same grid, keep the fastest for this session and cache it per grid
 * (padded dims, CPML, boundaries) for the process.  Candidates: the direct
 * family; 64x4 and 64x8 tiles with z-chunks of 8-32 planes and the lengths that
 * fill whole rounds of resident workgroups; and the balanced one-round launch
 * (tile_y 8, z_chunk 0; see seis3d_set_tiles). */
int seis3d_autotune(seis3d_t* t, const seis3d_source* src, int steps, int* family, int* tile_y,
                    int* z_chunk, double* best_ms_per_step);

typedef struct seis3d_batch seis3d_batch_t;
/* Capture n_steps steps of every session (each with its source, receivers and
 * an open gather) into one HIP graph: each shot runs on its own stream, the
 * graph is instantiated once and replayed with one launch.  Replaces the
 * sequential shot loop of the reference (src/Seis3D_Sim.cpp:133-143,
 * src/Seis3D/Acquisition3D.cpp:233-242).
 * The graph holds each session's device buffers.  While a batch exists, calls
 * that would free or replace them (seis3d_refresh_cpml, seis3d_upload_plane,
 * seis3d_finish_upload, seis3d_set_receivers, seis3d_gather_begin/_end,
 * seis3d_abc_attach/_detach) return SEIS3D_ELIFECYCLE, and seis3d_destroy of a
 * batched session is deferred until its last batch is destroyed.  Samples are
 * counted into each session's gather when the batch is launched; every launch
 * runs steps 0..n_steps-1 from the sessions' current wavefields and writes the
 * same gather slots. */
int seis3d_batch_create(seis3d_t* const* sessions, const seis3d_source* sources, int count,
                        int64_t n_steps, seis3d_batch_t** out);
/* Batch modes.  SHOT_STREAMS (what seis3d_batch_create builds): every shot's own launches on its
 * own stream.  SHARED_LAUNCHES: shots grouped by tile height, up to 4 per group, each group one
 * launch per half-step for all its shots (the shots' tiles dealt out as one list) on its own
 * stream.  WORK_QUEUE: one persistent launch per half-step for every shot of the batch: a grid of
 * resident 64x4 workgroups pulls (shot, tile, z-chunk) items from per-XCD queues, so shots of
 * uneven size share the GPU without a tail per shot; each shot is stepped in 64x4 tiles with its own
 * (autotuned) z-chunk length.
 * Shots with reflecting walls or a learned frame make the last two fall back to SHOT_STREAMS.
 * Every mode gives every shot's gather bit for bit as the shot run alone. */
enum seis3d_batch_mode {
    SEIS3D_BATCH_SHOT_STREAMS = 0,
    SEIS3D_BATCH_SHARED_LAUNCHES = 1,
    SEIS3D_BATCH_WORK_QUEUE = 2
};
int seis3d_batch_create_mode(seis3d_t* const* sessions, const seis3d_source* sources, int count,
                             int64_t n_steps, int mode, seis3d_batch_t** out);
int seis3d_batch_launch(seis3d_batch_t* b);
int seis3d_batch_synchronize(seis3d_batch_t* b);
void seis3d_batch_destroy(seis3d_batch_t* b);

/* ---- FWI gradient (no reference counterpart) -------------------------------- */

/* One shot of a gradient request: its source, its receivers (physical
 * coordinates, as seis3d_set_receivers) and the observed data, host memory
 * laid out as the gather: [n_steps / stride][receiver_count][vx, vy, vz]. */
typedef struct seis3d_shot {
    seis3d_source source;
    const seis3d_receiver* receivers;
    int32_t receiver_count;
    const float* observed;
} seis3d_shot;

/* Device bytes one shot of a gradient request needs, and the checkpoint
 * interval used (checkpoint_interval <= 0 picks ceil(sqrt(n_steps))). */
int seis3d_fwi_plan(const seis3d_ctx* ctx, int64_t n_steps, int stride, int checkpoint_interval,
                    int receivers, uint64_t* shot_bytes, int* interval);

/* Misfit 0.5 * sum over shots, samples, receivers and components of
 * (synthetic - observed)^2 for n_steps steps sampled after steps stride,
 * 2 stride, ... (as the gather), and its gradient with respect to vp, vs and
 * rho at every physical cell, [nz][ny][nx] host arrays, by the discrete
 * adjoint of the step.  The forward state is checkpointed every
 * checkpoint_interval steps and recomputed segment by segment.  Shots share
 * the session's model and run concurrently as device memory allows; the
 * session's own wavefield, receivers and gather are left untouched.
 * Any boundary: CPML, a free top, or reflecting walls. */
int seis3d_fwi_gradient(seis3d_t* t, const seis3d_shot* shots, int count, int64_t n_steps, int stride,
                        int checkpoint_interval, double* misfit, float* g_vp, float* g_vs, float* g_rho);

/* ---- Shot records and model files (host only) ------------------------------ */

/* Frozen acquisition metadata of one gather (Gather3D, include/Seis3D_Sim.h:47-57,
 * with Survey3D's world origin and CMP bin size, :36-45). */
typedef struct seis3d_gather_info {
    float spacing, frequency;          /* context the gather was recorded with */
    seis3d_source source;
    double world_origin[2];
    float cmp_bin[2];
    int32_t shot_id;
    int32_t stride;
    int32_t samples;
    int32_t sample_interval_us;
    int32_t coordinate_scalar;         /* SEG-Y XY scalar (-1000, -100, -10 or 1) */
    double first_sample_time;
} seis3d_gather_info;

/* Acquisition metadata of startGather3D (src/Seis3D/Acquisition3D.cpp:188-230):
 * validates source, sample count (1..32767), stride (1..10000), CMP bins and
 * world origin; interval = llround(dt 1e6) stride (<= 32767 us); first sample
 * time (stride - 1/2) dt; the largest XY scalar that keeps every coordinate
 * in 32-bit range. */
int seis3d_gather_info_init(const seis3d_ctx* ctx, const seis3d_source* src, const seis3d_receiver* receivers,
                            int count, int shot_id, int stride, int samples, const double world_origin[2],
                            const float cmp_bin[2], seis3d_gather_info* out);
/* exportGather3D (src/Seis3D/Acquisition3D.cpp:157-185): <prefix>_<shot:04>_vx.sgy,
 * _vy.sgy, _vz.sgy (SEG-Y rev1, IEEE float, big-endian; writeSegyComponent3D,
 * :76-154) and _geometry.csv in `directory`; refuses to overwrite any of them.
 * samples is the gather [sample][receiver][vx,vy,vz]. */
int seis3d_export_gather(const seis3d_gather_info* info, const seis3d_receiver* receivers, int count,
                         const float* samples, const char* directory, const char* prefix);

/* One imported single-component shot (readShotSegy3D, src/Seis3D/ShotRecordIo3D.cpp:66-141). */
typedef struct seis3d_segy_info {
    int32_t traces, samples, sample_interval_us, format, revision;
    int32_t component;                 /* 0 vx (id 14), 1 vy (13), 2 vz (12 or other) */
    int32_t shot_id;
    int32_t measurement_unit;          /* trace bytes 203-204 (6: m/s) */
    int32_t coordinate_units;          /* trace bytes 89-90 (1: length) */
} seis3d_segy_info;
/* Header scan and full validation of a rev0/rev1 big-endian file (formats 1, 2, 3, 5, 8). */
int seis3d_segy_inspect(const char* path, seis3d_segy_info* info);
/* Samples [trace][sample], receiver coordinates [trace][x, y, z] (metres, z as depth, positive down),
 * first-sample times [trace] and trace ids [trace]; any output may be null. */
int seis3d_segy_read(const char* path, float* samples, double* coordinates, double* first_times, int32_t* trace_ids);

/* inspectModel3D (src/Seis3D/ModelIo3D.cpp:35-65): header, payload size and every
 * material cell of a .s3dm file; fills nx, ny, nz, pml, boundaries, spacing,
 * max_vp and validates.  Other members of ctx are kept. */
int seis3d_s3dm_inspect(const char* path, seis3d_ctx* ctx);
/* importModel3D (:68-77): stream the planes into a session of the same shape. */
int seis3d_s3dm_load(seis3d_t* t, const char* path);
/* exportModel3D (:80-102): a new file only, written through <path>.part. */
int seis3d_s3dm_save(seis3d_t* t, const char* path);
/* The same file from host planes [z][y][x][vp, vs, rho] (no device needed). */
int seis3d_s3dm_write(const char* path, const seis3d_ctx* ctx, const float* vp_vs_rho);

/* ---- Learned absorbing boundary (no reference counterpart) ---------------- */

/* Number of FP32 parameters of the frame MLP: W1[96][32], b1[32], W2[32][32],
 * b2[32], W3[32][9], b3[9] (row-major, input index first). */
int seis3d_abc_param_count(void);
/* Replace the 8 CPML cells of every absorbing face (pml must be 8; a free top
 * keeps its free surface) by the learned frame: the frame takes the plain
 * update and, when params is not null, the MLP's correction after every
 * step (BF16 weights on MFMA).  params null leaves the plain update only
 * (the training base).  Calls again replace the parameters. */
int seis3d_abc_attach(seis3d_t* t, const float* params);
/* Restore the session's CPML: plain 8-cell CPML is the fallback. */
int seis3d_abc_detach(seis3d_t* t);
int seis3d_abc_frame_cells(const seis3d_t* t, int* count);
/* Copy the 9 wavefields of dst's non-ghost cells from src at cell offset
 * (ox, oy, oz) of src's padded grid; dst's ghost cells are cleared. */
int seis3d_copy_wavefield_box(seis3d_t* dst, const seis3d_t* src, int ox, int oy, int oz);

/* GPU trainer of the frame MLP.  Samples are teacher-forced: `thin` (learned
 * frame attached without parameters) holds one plain step taken from the
 * state of `wide` before `wide` advanced one step; the targets are wide's
 * values at the frame cells (thin cell + (ox, oy, oz)) minus thin's, in the
 * MLP's normalised face frame.  Up to max_cells random frame cells whose
 * field scale exceeds min_rel_scale times the largest are added. */
typedef struct seis3d_abc_trainer seis3d_abc_trainer_t;
int seis3d_abc_trainer_create(int capacity, uint64_t seed, seis3d_abc_trainer_t** out);
void seis3d_abc_trainer_destroy(seis3d_abc_trainer_t* tr);
int seis3d_abc_trainer_count(const seis3d_abc_trainer_t* tr);
int seis3d_abc_trainer_add(seis3d_abc_trainer_t* tr, seis3d_t* thin, const seis3d_t* wide, int ox, int oy, int oz,
                           int max_cells, float min_rel_scale, int* added);
/* Adam over shuffled mini-batches (cosine-decayed rate); params in/out; mean squared error per epoch. */
int seis3d_abc_train(seis3d_abc_trainer_t* tr, float* params, int epochs, int batch, float lr, float* epoch_loss);
int seis3d_abc_evaluate(seis3d_abc_trainer_t* tr, const float* params, double* loss);
/* The training kernels on host data: forward (Y [n][9], BF16-rounded H1/H2
 * [n][32]), mean squared error against T and the parameter gradient. */
int seis3d_abc_mlp(const float* params, const float* X, const float* T, int n, float* Y, float* H1, float* H2,
                   float* grad, double* loss);

/* ---- Errors and build info ------------------------------------------------ */

/* HBM bandwidth probe (no reference counterpart; the measured ceiling the step kernels are
 * compared with).  Streams arrays of `bytes` each, `reps` timed launches after one warm-up:
 * kind 0 read, 1 write, 2 copy, 3 copy staged through LDS by global_load_lds, 4 ten inputs and
 * three outputs elementwise, 5 twelve inputs and six outputs.  *gbs = bytes moved per second
 * (every array touched once per launch), *us_per_launch the average launch time. */
int seis3d_bandwidth_probe(int kind, uint64_t bytes, int reps, double* gbs, double* us_per_launch);

const char* seis3d_last_error(void);
const char* seis3d_version(void);
/* Number of visible HIP devices (0 when none or the runtime is unusable). */
int seis3d_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* SEIS3D_H */
