// Seis2D.hpp - header-only C++ layer over the 2D C ABI (seis2d.h) with the
// reference's 2D API names and types (CUDA_Kernels/Cuda_Check.cuh:127-149,
// Main_header/Common.h:251-287, Main_header/SolverNumerics.h).
//
// Callers of the reference keep their code: SimulationContext has the same
// members (std::vector arrays), the functions keep their names, argument
// meaning and exceptions (std::invalid_argument for bad input,
// std::runtime_error for device failures, std::overflow_error for the index
// capacity check).  GPUSimData owns an opaque session handle instead of raw
// device pointers; the reference session's direct copy into d_active_sources becomes
// uploadActiveSources().
#pragma once

#include <array>
#include <cstddef>
#include <filesystem>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "seis2d.h"

constexpr int kStencilRadius = 4;
constexpr int kMaximumActiveSources = 200;
constexpr float kSourceInjectionScale2D = 1e7f;
constexpr std::array<double, 4> kStaggeredDerivativeWeights = {1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0,
                                                               -5.0 / 7168.0};

struct GPUSource {
    int idx;
    float t;
    float f_peak;
    float amp;
};

struct SimulationContext {
    int NX = 0, NZ = 0;
    int total_grid = 0;
    float h = 0.0f, dt = 0.0f;
    int nt = 0, npml = 0;
    int snap_spacenum = 10;
    int flag_type = 1;
    std::vector<float> rho, mu, lambda, lambda2mu;
    std::vector<float> dp_flat;
    std::vector<float> dx, dx_half, dz, dz_half;
    float c1_h = 0.0f, c2_h = 0.0f, c3_h = 0.0f, c4_h = 0.0f;
    int src_idx = 0;
    int src_z_idx = 0;
    std::vector<float> wavelet;
    float src_angle = 0.0f;
    int src_type = 0;
    float src_amp = 1.0f;
    int num_rcv = 0;
    std::vector<int> rcv_grid_idx;
    std::vector<float> record_vx;
    std::vector<float> record_vz;
    bool upFlag = false;
};

struct GPUSimData {
    seis2d_t* session = nullptr;
    int NX = 0, NZ = 0, total_grid = 0, flag_type = 1, receiverCapacity = 0;
};

namespace seis2d_detail {

inline void check(int rc) {
    if (rc == SEIS2D_OK) return;
    const std::string msg = seis2d_last_error();
    if (rc == SEIS2D_EINVAL) throw std::invalid_argument(msg);
    if (rc == SEIS2D_ELIFECYCLE) throw std::logic_error(msg);
    throw std::runtime_error(msg);
}

inline seis2d_ctx scalars(const SimulationContext& c) {
    seis2d_ctx o{};
    o.nx = c.NX; o.nz = c.NZ; o.total_grid = c.total_grid; o.h = c.h; o.dt = c.dt; o.nt = c.nt; o.npml = c.npml;
    o.snap_spacenum = c.snap_spacenum; o.flag_type = c.flag_type;
    o.c1_h = c.c1_h; o.c2_h = c.c2_h; o.c3_h = c.c3_h; o.c4_h = c.c4_h;
    o.src_idx = c.src_idx; o.src_z_idx = c.src_z_idx; o.src_angle = c.src_angle; o.src_type = c.src_type;
    o.src_amp = c.src_amp; o.num_rcv = c.num_rcv; o.up_flag = c.upFlag ? 1 : 0;
    return o;
}

// Array lengths are checked here, as the reference checks std::vector sizes (Cuda_Check.cu:850-857).
inline seis2d_arrays arrays(const SimulationContext& c) {
    const std::size_t n = std::size_t(c.NX > 0 ? c.NX : 0) * std::size_t(c.NZ > 0 ? c.NZ : 0);
    if (c.rho.size() != n || c.mu.size() != n || c.lambda.size() != n || c.lambda2mu.size() != n ||
        c.dx.size() != std::size_t(c.NX) || c.dx_half.size() != std::size_t(c.NX) || c.dz.size() != std::size_t(c.NZ) ||
        c.dz_half.size() != std::size_t(c.NZ) || (c.flag_type == 3 && c.dp_flat.size() != std::size_t(c.NX)) ||
        c.wavelet.size() < std::size_t(c.nt > 0 ? c.nt : 0) || c.num_rcv < 0 ||
        c.rcv_grid_idx.size() != std::size_t(c.num_rcv))
        throw std::invalid_argument("Simulation array lengths do not match the grid or receiver count");
    seis2d_arrays a{};
    a.rho = c.rho.data(); a.mu = c.mu.data(); a.lambda = c.lambda.data(); a.lambda2mu = c.lambda2mu.data();
    a.dp_flat = c.dp_flat.empty() ? nullptr : c.dp_flat.data();
    a.dx = c.dx.data(); a.dx_half = c.dx_half.data(); a.dz = c.dz.data(); a.dz_half = c.dz_half.data();
    a.wavelet = c.wavelet.data(); a.wavelet_len = int32_t(c.wavelet.size());
    a.rcv_grid_idx = c.rcv_grid_idx.empty() ? nullptr : c.rcv_grid_idx.data();
    return a;
}

}  // namespace seis2d_detail

// checkedGridElementCount (SolverNumerics.h:25-34).
inline std::size_t checkedGridElementCount(int nx, int nz) {
    int64_t n = 0;
    const int rc = seis2d_grid_count(nx, nz, &n);
    if (rc == SEIS2D_EINVAL && nx > 2 * kStencilRadius && nz > 2 * kStencilRadius)
        throw std::overflow_error(seis2d_last_error());
    seis2d_detail::check(rc);
    return std::size_t(n);
}
// validateBoundaryGeometry (SolverNumerics.h:37-46).
inline void validateBoundaryGeometry(int nx, int nz, int pml, bool freeSurface) {
    seis2d_detail::check(seis2d_boundary_geometry(nx, nz, pml, freeSurface ? 1 : 0));
}
// getReceiverSamplingStride (SolverNumerics.h:49-59).
inline int getReceiverSamplingStride(double dt, double rateHz) {
    int s = 1;
    seis2d_detail::check(seis2d_sampling_stride(dt, rateHz, &s));
    return s;
}

// generateRickerWavelet (src/Seis2D/Numerics2D.inc:3-14).
inline void generateRickerWavelet(std::vector<float>& wavelet, int nt, float dt, float f0, float t0) {
    wavelet.resize(std::size_t(nt > 0 ? nt : 0));
    seis2d_detail::check(seis2d_ricker_wavelet(nt, dt, f0, t0, wavelet.data()));
}

// updatePmlDampingArrays (Numerics2D.inc:36-83) on a context.
inline void updatePmlDampingArrays(SimulationContext& c) {
    c.dx.assign(std::size_t(c.NX), 0.0f); c.dx_half.assign(std::size_t(c.NX), 0.0f);
    c.dz.assign(std::size_t(c.NZ), 0.0f); c.dz_half.assign(std::size_t(c.NZ), 0.0f);
    const seis2d_ctx s = seis2d_detail::scalars(c);
    seis2d_detail::check(seis2d_pml_profiles(&s, c.lambda2mu.data(), c.rho.data(), c.dx.data(), c.dx_half.data(),
                                             c.dz.data(), c.dz_half.data()));
}

// validateSimulationContext (Cuda_Check.cu:841-884).
inline void validateSimulationContext(const SimulationContext& c) {
    const seis2d_arrays a = seis2d_detail::arrays(c);
    const seis2d_ctx s = seis2d_detail::scalars(c);
    seis2d_detail::check(seis2d_validate(&s, &a));
}

// releaseGpuSimulation (Cuda_Check.cu:970-990): safe to repeat.
inline void releaseGpuSimulation(GPUSimData& g) {
    seis2d_destroy(g.session);
    g = GPUSimData{};
}

// initializeGpuSimulation (Cuda_Check.cu:887-967).
inline void initializeGpuSimulation(GPUSimData& g, const SimulationContext& c) {
    const seis2d_arrays a = seis2d_detail::arrays(c);
    const seis2d_ctx s = seis2d_detail::scalars(c);
    releaseGpuSimulation(g);
    seis2d_detail::check(seis2d_create(&s, &a, &g.session));
    g.NX = c.NX; g.NZ = c.NZ; g.total_grid = c.total_grid; g.flag_type = c.flag_type; g.receiverCapacity = c.num_rcv;
}

// updateGpuReceivers (Cuda_Check.cu:992-1024).
inline void updateGpuReceivers(GPUSimData& g, const std::vector<int>& idx) {
    seis2d_detail::check(seis2d_update_receivers(g.session, idx.empty() ? nullptr : idx.data(), int(idx.size())));
    g.receiverCapacity = int(idx.size());
}

// The 2D driver's copy into d_active_sources (src/Seis2D/Workflow2D.inc:313-315).
inline void uploadActiveSources(GPUSimData& g, const std::vector<GPUSource>& sources) {
    std::vector<seis2d_source> s(sources.size());
    for (std::size_t k = 0; k < sources.size(); ++k) s[k] = {sources[k].idx, sources[k].t, sources[k].f_peak, sources[k].amp};
    seis2d_detail::check(seis2d_set_sources(g.session, s.data(), int(s.size())));
}

// runGpuSimulationStep (Cuda_Check.cu:1027-1085).
inline void runGpuSimulationStep(GPUSimData& g, int it, const SimulationContext& c, int activeSources) {
    const seis2d_ctx s = seis2d_detail::scalars(c);
    seis2d_detail::check(seis2d_step(g.session, it, &s, c.wavelet.empty() ? nullptr : c.wavelet.data(),
                                     int(c.wavelet.size()), activeSources));
}

// copyWavefieldToHost (Cuda_Check.cu:1088-1103).
inline void copyWavefieldToHost(GPUSimData& g, float* vz) { seis2d_detail::check(seis2d_copy_wavefield(g.session, vz)); }

// copyReceiverRecordsFromGpu (Cuda_Check.cu:1106-1134).
inline void copyReceiverRecordsFromGpu(GPUSimData& g, int numRcv, int nt, int it, float* recVx, float* recVz) {
    seis2d_detail::check(seis2d_copy_receiver_records(g.session, numRcv, nt, it, recVx, recVz));
}

// recordReceiverStepOnGpu (Cuda_Check.cu:1147-1169).
inline void recordReceiverStepOnGpu(GPUSimData& g, int numRcv, float* vx, float* vz) {
    seis2d_detail::check(seis2d_record_receiver_step(g.session, numRcv, vx, vz));
}

// ---- shot records and shot packages ----

namespace SeismicIO {

// TraceMetadata (Main_header/SeismicIO.h:21-29).
struct TraceMetadata {
    int32_t fieldRecordNum = 1;
    int32_t traceInRecord = 1;
    int32_t sourceX = 0, sourceY = 0;
    int32_t groupX = 0, groupY = 0;
    int16_t coordinateScalar = 1;
};

inline std::vector<seis2d_trace_meta> toC(const std::vector<TraceMetadata>& m) {
    std::vector<seis2d_trace_meta> o(m.size());
    for (std::size_t i = 0; i < m.size(); ++i)
        o[i] = {m[i].fieldRecordNum, m[i].traceInRecord, m[i].sourceX, m[i].sourceY, m[i].groupX, m[i].groupY,
                m[i].coordinateScalar};
    return o;
}

inline std::vector<float> flatten(const std::vector<std::vector<float>>& traces, std::size_t samples) {
    std::vector<float> flat;
    flat.reserve(traces.size() * samples);
    for (const auto& t : traces) {
        if (t.size() != samples) throw std::invalid_argument("SEG-Y export traces have different lengths.");
        flat.insert(flat.end(), t.begin(), t.end());
    }
    return flat;
}

// writeSegyCaptureFile2D (Main_cpp/SeismicIO.cpp:296-414).
inline void writeSegyCaptureFile2D(const std::vector<std::vector<float>>& data, const std::vector<TraceMetadata>& meta,
                                   const std::filesystem::path& path, float dt) {
    if (data.empty() || data.front().empty()) throw std::invalid_argument("SEG-Y export requires samples and a positive interval.");
    if (meta.size() != data.size())
        throw std::invalid_argument("SEG-Y sample count/interval must fit 1..32767 and geometry must match.");
    const auto flat = flatten(data, data.front().size());
    const auto m = toC(meta);
    seis2d_detail::check(seis2d_write_segy(path.string().c_str(), flat.data(), int(data.size()),
                                           int(data.front().size()), m.data(), dt));
}

}  // namespace SeismicIO

namespace Seis2D {

// RecordData2D of readRecordSegy2D (src/Seis2D/RecordData2D.cpp:35-65).
struct RecordData2D {
    std::vector<std::vector<float>> traces;
    std::vector<double> firstTimes;
    double sampleInterval = 0;
    std::string name, units;
};

inline RecordData2D readRecordSegy2D(const std::filesystem::path& path) {
    seis2d_record_info info{};
    seis2d_detail::check(seis2d_read_segy_info(path.string().c_str(), &info));
    std::vector<float> flat(std::size_t(info.traces) * std::size_t(info.samples));
    RecordData2D r;
    r.firstTimes.resize(std::size_t(info.traces));
    seis2d_detail::check(seis2d_read_segy(path.string().c_str(), flat.data(), r.firstTimes.data()));
    r.traces.resize(std::size_t(info.traces));
    for (std::size_t c = 0; c < r.traces.size(); ++c)
        r.traces[c].assign(flat.begin() + std::ptrdiff_t(c * std::size_t(info.samples)),
                           flat.begin() + std::ptrdiff_t((c + 1) * std::size_t(info.samples)));
    r.sampleInterval = info.sample_interval_us * 1e-6;
    r.name = path.filename().string();
    if (info.velocity_units) r.units = "m/s";
    return r;
}

// FrozenModel2D and ShotExportRequest2D (Main_header/ShotExport2D.h:7-19).
struct FrozenModel2D {
    int nx = 0, nz = 0, npml = 0, formulation = 0, sourceX = 0, sourceZ = 0, sourceType = 0, captureStartStep = 0,
        samplingStride = 1;
    bool freeSurface = false, multipleSources = false;
    float spacing = 0, solverDt = 0, frequency = 0, delay = 0, sourceAmplitude = 0, sourceAngle = 0;
    std::vector<float> lambda, mu, rho, wavelet;
};
struct ShotExportRequest2D {
    std::filesystem::path directory;
    std::string prefix = "seis2d_shot", version, device;
    int shot = 1, samples = 0;
    float sampleInterval = 0;
    bool includeCsv = true, includeModel = true;
};

// validateExportPrefix2D (src/Seis2D/ShotExport2D.cpp:103-114).
inline void validateExportPrefix2D(const std::string& prefix) {
    seis2d_detail::check(seis2d_validate_prefix(prefix.c_str()));
}

// exportShotPackage2D (src/Seis2D/ShotExport2D.cpp:117-187): the published folder.
inline std::filesystem::path exportShotPackage2D(const ShotExportRequest2D& r, const FrozenModel2D& m,
                                                 const std::vector<std::vector<float>>& vx,
                                                 const std::vector<std::vector<float>>& vz,
                                                 const std::vector<SeismicIO::TraceMetadata>& geometry) {
    if (vx.empty() || vx.size() != vz.size() || vx.size() != geometry.size())
        throw std::invalid_argument("Shot components and geometry differ.");
    for (const auto* c : {&vx, &vz})
        for (const auto& t : *c)
            if (t.size() != std::size_t(r.samples)) throw std::invalid_argument("Shot sample counts differ.");
    const auto fx = SeismicIO::flatten(vx, std::size_t(r.samples)), fz = SeismicIO::flatten(vz, std::size_t(r.samples));
    const auto g = SeismicIO::toC(geometry);
    const std::size_t cells = std::size_t(m.nx > 0 ? m.nx : 0) * std::size_t(m.nz > 0 ? m.nz : 0);
    if (r.includeModel && (m.lambda.size() != cells || m.mu.size() != cells || m.rho.size() != cells))
        throw std::invalid_argument("Frozen model arrays are absent, inconsistent or exceed 256 MiB.");
    seis2d_frozen_model cm{m.nx, m.nz, m.npml, m.formulation, m.sourceX, m.sourceZ, m.sourceType, m.captureStartStep,
                           m.samplingStride, m.freeSurface, m.multipleSources, m.spacing, m.solverDt, m.frequency,
                           m.delay, m.sourceAmplitude, m.sourceAngle, m.lambda.data(), m.mu.data(), m.rho.data(),
                           m.wavelet.data(), int32_t(m.wavelet.size())};
    const std::string dir = r.directory.string();
    seis2d_export_request cr{dir.c_str(), r.prefix.c_str(), r.version.c_str(), r.device.c_str(), r.shot, r.samples,
                             r.sampleInterval, r.includeCsv, r.includeModel};
    std::vector<char> out(4096);
    seis2d_detail::check(seis2d_export_shot_package(&cr, &cm, fx.data(), fz.data(), int(vx.size()), g.data(),
                                                    out.data(), out.size()));
    return std::filesystem::path(out.data());
}

}  // namespace Seis2D
