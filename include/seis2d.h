/* seis2d.h - C ABI of the MI355X 2D P-SV velocity-stress solver.
 *
 * Replaces the reference's 2D solver API (CUDA_Kernels/Cuda_Check.cuh:127-149)
 * on the data of SimulationContext (Main_header/Common.h:259-287), with the
 * host numerics of Main_header/SolverNumerics.h and src/Seis2D/Numerics2D.inc.
 * The reference takes std::vector members; here the arrays are plain pointers
 * with the same meaning and lengths.  Grids are row-major [z][x] with flat
 * index k = z * nx + x, as the reference.
 *
 * Every call returns 0 on success or an error code; seis2d_last_error()
 * gives the message (thread-local).  Codes: 1 invalid argument
 * (std::invalid_argument in the reference), 2 HIP failure (the reference's
 * CUDA_CHECK runtime_error), 3 lifecycle misuse, 4 file runtime error
 * (std::runtime_error: unreadable input, existing destination).
 */
#ifndef SEIS2D_H
#define SEIS2D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SEIS2D_OK = 0, SEIS2D_EINVAL = 1, SEIS2D_EHIP = 2, SEIS2D_ELIFECYCLE = 3, SEIS2D_EIO = 4 };

/* Boundary schemes (SimulationContext::flag_type). */
enum { SEIS2D_SPLIT_PML = 1, SEIS2D_HYBRID_PML = 2, SEIS2D_FREE_SURFACE = 3 };

/* Source kinds (GPUSimData::src_type, Cuda_Check.cu:574-579). */
enum {
    SEIS2D_VERTICAL_FORCE = 0,
    SEIS2D_HORIZONTAL_FORCE = 1,
    SEIS2D_EXPLOSIVE = 2,
    SEIS2D_SHEAR = 3,
    SEIS2D_ROTATED_FORCE = 4
};

/* Field identifiers for seis2d_download_field / seis2d_upload_field.
 * Split pairs (types 1, 3 and the type-2 PML frame), then type 2's combined fields. */
enum {
    SEIS2D_VX1 = 0, SEIS2D_VX2 = 1, SEIS2D_VZ1 = 2, SEIS2D_VZ2 = 3,
    SEIS2D_SXX1 = 4, SEIS2D_SXX2 = 5, SEIS2D_SZZ1 = 6, SEIS2D_SZZ2 = 7,
    SEIS2D_SXZ1 = 8, SEIS2D_SXZ2 = 9,
    SEIS2D_VX = 10, SEIS2D_VZ = 11, SEIS2D_SXX = 12, SEIS2D_SZZ = 13, SEIS2D_SXZ = 14
};

/* Scalars of SimulationContext (Main_header/Common.h:259-287). */
typedef struct seis2d_ctx {
    int32_t nx, nz, total_grid;
    float h, dt;
    int32_t nt, npml, snap_spacenum;
    int32_t flag_type;                  /* 1 split PML, 2 hybrid PML/FDM, 3 free surface */
    float c1_h, c2_h, c3_h, c4_h;       /* O(8) weights / h */
    int32_t src_idx, src_z_idx;         /* flat source index and its row */
    float src_angle;                    /* degrees, rotated force */
    int32_t src_type;
    float src_amp;
    int32_t num_rcv;
    int32_t up_flag;                    /* free surface on (type 3) */
} seis2d_ctx;

/* Array members of SimulationContext; lengths as the reference validates them
 * (validateSimulationContext, Cuda_Check.cu:841-884). */
typedef struct seis2d_arrays {
    const float* rho;                   /* nx * nz */
    const float* mu;
    const float* lambda;
    const float* lambda2mu;
    const float* dp_flat;               /* nx, free-surface modulus (type 3 only) */
    const float* dx;                    /* nx whole-grid PML damping */
    const float* dx_half;               /* nx */
    const float* dz;                    /* nz */
    const float* dz_half;               /* nz */
    const float* wavelet;               /* >= nt samples */
    int32_t wavelet_len;
    const int32_t* rcv_grid_idx;        /* num_rcv flat indices */
} seis2d_arrays;

/* GPUSource (Main_header/Common.h:251-256): one active point source. */
typedef struct seis2d_source {
    int32_t idx;                        /* flat index */
    float t;                            /* elapsed time of the source */
    float f_peak;
    float amp;
} seis2d_source;

typedef struct seis2d_session seis2d_t;

/* ---- Host numerics (no device needed) ------------------------------------ */

/* kCourantLimit2d (Main_header/SolverNumerics.h:18-22). */
float seis2d_courant_limit(void);
/* checkedGridElementCount (SolverNumerics.h:25-34). */
int seis2d_grid_count(int nx, int nz, int64_t* count);
/* validateBoundaryGeometry (SolverNumerics.h:37-46). */
int seis2d_boundary_geometry(int nx, int nz, int npml, int free_surface);
/* getReceiverSamplingStride (SolverNumerics.h:49-59). */
int seis2d_sampling_stride(double dt, double rate_hz, int* stride);
/* generateRickerWavelet (src/Seis2D/Numerics2D.inc:3-14). */
int seis2d_ricker_wavelet(int nt, float dt, float f0, float t0, float* out);
/* updatePmlDampingArrays (Numerics2D.inc:36-83): max vp over the model
 * (<= 350 m/s falls back to 2000), then the power-2.5 profiles. */
int seis2d_pml_profiles(const seis2d_ctx* ctx, const float* lambda2mu, const float* rho, float* dx, float* dx_half,
                        float* dz, float* dz_half);
/* validateSimulationContext (Cuda_Check.cu:841-884). */
int seis2d_validate(const seis2d_ctx* ctx, const seis2d_arrays* arrays);

/* ---- Session ----------------------------------------------------------- */

/* initializeGpuSimulation (Cuda_Check.cu:887-967): validate, allocate, upload. */
int seis2d_create(const seis2d_ctx* ctx, const seis2d_arrays* arrays, seis2d_t** out);
/* releaseGpuSimulation (Cuda_Check.cu:970-990); null is a no-op. */
void seis2d_destroy(seis2d_t* t);
/* updateGpuReceivers (Cuda_Check.cu:992-1024): atomic replace; count 0 clears. */
int seis2d_update_receivers(seis2d_t* t, const int32_t* idx, int count);
/* The active-source upload of the 2D driver (src/Seis2D/Workflow2D.inc:298-315), at most 200. */
int seis2d_set_sources(seis2d_t* t, const seis2d_source* sources, int count);
/* runGpuSimulationStep (Cuda_Check.cu:1027-1085).  The scalars of ctx that the
 * reference reads each step (src_type, src_angle, src_amp, src_idx, src_z_idx,
 * nt) come from ctx; the single-source sample is wavelet[it] (wavelet_len
 * samples).  active_sources >= 0 injects that many uploaded sources instead. */
int seis2d_step(seis2d_t* t, int it, const seis2d_ctx* ctx, const float* wavelet, int wavelet_len, int active_sources);
/* copyWavefieldToHost (Cuda_Check.cu:1088-1103): vz (split pairs summed). */
int seis2d_copy_wavefield(seis2d_t* t, float* vz_out);
/* copyReceiverRecordsFromGpu (Cuda_Check.cu:1106-1134): traces [r][nt], column it. */
int seis2d_copy_receiver_records(seis2d_t* t, int num_rcv, int nt, int it, float* rec_vx, float* rec_vz);
/* recordReceiverStepOnGpu (Cuda_Check.cu:1147-1169). */
int seis2d_record_receiver_step(seis2d_t* t, int num_rcv, float* vx, float* vz);
int seis2d_synchronize(seis2d_t* t);

/* ---- No reference counterpart ------------------------------------------- */

/* n_steps single-source steps from first_it, replayed from one captured HIP
 * graph, with every receiver recorded after every step into device traces;
 * rec_vx / rec_vz (may be null) receive [num_rcv][n_steps] at the end.  Same
 * arithmetic as seis2d_step + seis2d_record_receiver_step per step. */
int seis2d_run(seis2d_t* t, const seis2d_ctx* ctx, const float* wavelet, int wavelet_len, int first_it, int n_steps,
               float* rec_vx, float* rec_vz);
/* One field (enum above) as [nz][nx]; unallocated fields give EINVAL. */
int seis2d_download_field(seis2d_t* t, int field, float* out);
int seis2d_upload_field(seis2d_t* t, int field, const float* in);

const char* seis2d_last_error(void);

/* ---- Shot records and shot packages (host only) ------------------------------ */

/* SeismicIO::TraceMetadata (Main_header/SeismicIO.h:21-29): trace header geometry. */
typedef struct seis2d_trace_meta {
    int32_t field_record;       /* bytes 9-12 */
    int32_t trace_in_record;    /* bytes 13-16 */
    int32_t source_x, source_y; /* bytes 73-80 (Y holds Z, down) */
    int32_t group_x, group_y;   /* bytes 81-88 */
    int32_t coordinate_scalar;  /* bytes 71-72 (int16 on disk) */
} seis2d_trace_meta;

/* writeSegyCaptureFile2D (Main_cpp/SeismicIO.cpp:296-414): one component, traces
 * [trace][sample], IEEE big-endian, m/s; every input is validated before the
 * file is opened. */
int seis2d_write_segy(const char* path, const float* traces, int n_traces, int n_samples,
                      const seis2d_trace_meta* meta, float dt);

/* readRecordSegy2D (src/Seis2D/RecordData2D.cpp:35-65). */
typedef struct seis2d_record_info {
    int32_t traces, samples, sample_interval_us, format;
    int32_t velocity_units;     /* every trace has measurement unit 6 (m/s) */
} seis2d_record_info;
int seis2d_read_segy_info(const char* path, seis2d_record_info* info);
/* traces [trace][sample]; first_times [trace] in seconds; either may be null. */
int seis2d_read_segy(const char* path, float* traces, double* first_times);

/* FrozenModel2D (Main_header/ShotExport2D.h:7-12): the solver inputs of a shot. */
typedef struct seis2d_frozen_model {
    int32_t nx, nz, npml, formulation, source_x, source_z, source_type, capture_start_step, sampling_stride;
    int32_t free_surface, multiple_sources;
    float spacing, solver_dt, frequency, delay, source_amplitude, source_angle;
    const float *lambda, *mu, *rho;   /* [nz][nx] */
    const float* wavelet;
    int32_t wavelet_count;
} seis2d_frozen_model;

/* ShotExportRequest2D (Main_header/ShotExport2D.h:13-19). */
typedef struct seis2d_export_request {
    const char* directory;
    const char* prefix;          /* default "seis2d_shot" */
    const char* version;
    const char* device;
    int32_t shot, samples;
    float sample_interval;
    int32_t include_csv, include_model;
} seis2d_export_request;

/* validateExportPrefix2D (src/Seis2D/ShotExport2D.cpp:103-114). */
int seis2d_validate_prefix(const char* prefix);

/* exportShotPackage2D (src/Seis2D/ShotExport2D.cpp:117-187): a new folder
 * <directory>/<prefix>_<shot:04> holding record_vx.sgy, record_vz.sgy,
 * geometry.csv, optionally record_vz.csv, model.s2dm and source_wavelet.f32le,
 * and manifest.json (schema SeisWaveShot2D-1) with each file's size and
 * SHA-256.  Written in a staging folder and renamed; an existing destination
 * is an error (code 4).  vx, vz are [trace][sample]; the folder path is copied
 * to out_path. */
int seis2d_export_shot_package(const seis2d_export_request* request, const seis2d_frozen_model* model,
                               const float* vx, const float* vz, int traces, const seis2d_trace_meta* geometry,
                               char* out_path, size_t out_capacity);

/* SHA-256 (FIPS 180-4) of a file, as 64 lowercase hex digits plus NUL. */
int seis2d_sha256_file(const char* path, char out_hex[65]);

#ifdef __cplusplus
}
#endif

#endif /* SEIS2D_H */
