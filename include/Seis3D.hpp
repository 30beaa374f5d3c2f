// Seis3D.hpp - C++ solver-session API of the reference (cuda/Cuda_Seis3D.cuh:6-42 on the POD
// types of include/Common3D.h:9-98), header-only, over the C ABI of seis3d.h.
//
// Callers of the reference keep their code: the same namespace, type and function names, the
// same field names and defaults, the same argument meaning, and the same exception types with
// the reference's message prefixes (std::invalid_argument for bad input, std::runtime_error for
// device failures and non-finite wavefields, std::logic_error for lifecycle misuse).  What
// differs is what the structs hold: GPUSimData3D owns an opaque session handle instead of raw
// device pointers, since the device layout (pitched rows, x-fastest CPML slabs) is the
// library's own.  Link with libseis3d.so.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "seis3d.h"

namespace Seis3D {

constexpr int stencilRadius3D = 4;
constexpr std::size_t maxPhysicalNodes3D = std::size_t(300) * 300 * 400;
constexpr int maxVolumeAxis3D = 256;
constexpr int maxVolumeOutput3D = 1536;
constexpr std::size_t memoryLimit3D = std::size_t(4) * 1024 * 1024 * 1024;
constexpr int maxReceivers3D = 300;
constexpr int maxSegySamples3D = 32767;
constexpr int maxSegyIntervalUs3D = 32767;

enum class Field3D : int { Vx, Vy, Vz, Sxx, Syy, Szz, Sxy, Syz, Sxz };
enum class ModelKind3D : int { Uniform, TwoLayer, Sphere, Imported, ThreeLayer, DippingLayer,
    LowVelocityLayer, Fault, Gradient, SaltDome };
enum class SourceKind3D : int { Explosion, Force };

// include/Common3D.h:24-42, same names and defaults.
struct SimulationContext3D {
    int nx3D = 300, ny3D = 50, nz3D = 300;
    int pml3D = 12;
    bool freeSurface3D = true;
    bool reflecting3D = false;
    float spacing3D = 10.0f, cfl3D = 0.85f, dt3D = 0.0f;
    bool manualTimeStep3D = false;
    float requestedDtMs3D = 0.5f;
    float safeDt3D = 0.0f;
    float vp3D = 3000.0f, vs3D = 1700.0f, density3D = 2200.0f;
    float layerVp3D = 4200.0f, layerVs3D = 2400.0f, layerDensity3D = 2500.0f;
    float maxVp3D = 4200.0f, frequency3D = 20.0f;
    ModelKind3D modelKind3D = ModelKind3D::TwoLayer;
    float interfaceDepth3D = 0.5f, layerDip3D = 0.25f, layerThickness3D = 0.15f;
    float faultThrow3D = 0.18f, anomalyRadius3D = 0.187f;
    int gx3D = 0, gy3D = 0, gz3D = 0, ox3D = 0, oy3D = 0, oz3D = 0;
    std::size_t cells3D = 0;
};

struct Source3D {
    int x3D = 48, y3D = 48, z3D = 24;
    SourceKind3D kind3D = SourceKind3D::Explosion;
    float amplitude3D = 1.0e8f;
    float direction3D[3] = {0.0f, 0.0f, 1.0f};
};

struct Receiver3D { int x3D = 0, y3D = 0, z3D = 0; };

struct MemoryBudget3D {
    std::size_t fieldBytes3D = 0, cpmlBytes3D = 0, viewBytes3D = 0, totalBytes3D = 0;
};

// Survey3D (include/Seis3D_Sim.h:36-45): receiver and shot spreads, world origin, CMP bins.
struct Survey3D {
    int receiverCountX3D = 50, receiverCountY3D = 3;
    int receiverOrigin3D[3] = {27, 1, 0};
    int receiverSpacing3D[2] = {5, 23};
    int shotCountX3D = 3, shotCountY3D = 2;
    int shotOrigin3D[3] = {24, 24, 12};
    int shotSpacing3D[2] = {20, 20};
    double worldOrigin3D[2] = {0.0, 0.0};
    float cmpBinSize3D[2] = {25.0f, 25.0f};
};

// Gather3D (include/Seis3D_Sim.h:47-57): frozen geometry and [sample][receiver][vx,vy,vz] samples.
struct Gather3D {
    SimulationContext3D context3D;
    Source3D source3D;
    Survey3D survey3D;
    std::vector<Receiver3D> receivers3D;
    std::vector<float> samples3D;
    int sampleCount3D = 0, targetSamples3D = 1000, stride3D = 2, shotId3D = 1;
    int sampleIntervalUs3D = 0;
    int coordinateScalar3D = -1000;
    double firstSampleTime3D = 0.0;
};

// ImportedShot3D (include/Seis3D_Sim.h:60-68): one single-component shot read from SEG-Y.
struct ImportedShot3D {
    Gather3D gather3D;
    std::vector<std::array<double, 3>> coordinates3D;
    std::vector<double> firstTimes3D;
    std::vector<int> traceIds3D;
    int component3D = 2;
    bool cartesian3D = true;
    std::string componentName3D = "Amplitude", amplitudeUnit3D = "file units", coordinateUnit3D = "m", fileName3D;
};

// ViewLayout3D (include/Common3D.h:53-60).
struct ViewLayout3D {
    std::size_t sliceOffset3D[3]{};
    int sliceWidth3D[3]{}, sliceHeight3D[3]{};
    std::size_t volumeOffset3D = 0, totalSamples3D = 0;
    int volumeX3D = 0, volumeY3D = 0, volumeZ3D = 0;
    int domainSize3D[3]{}, domainOrigin3D[3]{};
};

// The caller owns the struct; the library owns the session behind it.
struct GPUSimData3D {
    seis3d_t* session3D = nullptr;
    int receiverCount3D = 0;
    std::size_t allocatedBytes3D = 0;
};

inline int activePml3D(const SimulationContext3D& c) { return c.reflecting3D ? 0 : c.pml3D; }

inline const char* boundaryName3D(const SimulationContext3D& c) {
    return c.reflecting3D ? "Reflecting slip walls" : c.freeSurface3D ? "Free top + CPML" : "All faces CPML";
}

inline std::size_t gridIndex3D(const SimulationContext3D& c, int x, int y, int z) {
    return (std::size_t(z) * c.gy3D + y) * c.gx3D + x;
}

namespace detail {

inline void check(int rc) {
    if (rc == SEIS3D_OK) return;
    const std::string message = seis3d_last_error();
    if (rc == SEIS3D_EINVAL) throw std::invalid_argument(message);
    if (rc == SEIS3D_ELIFECYCLE) throw std::logic_error(message);
    throw std::runtime_error(message);
}

inline seis3d_ctx toC(const SimulationContext3D& c) {
    seis3d_ctx o{};
    o.nx = c.nx3D; o.ny = c.ny3D; o.nz = c.nz3D; o.pml = c.pml3D;
    o.free_surface = c.freeSurface3D; o.reflecting = c.reflecting3D;
    o.spacing = c.spacing3D; o.cfl = c.cfl3D; o.dt = c.dt3D;
    o.manual_dt = c.manualTimeStep3D; o.requested_dt_ms = c.requestedDtMs3D; o.safe_dt = c.safeDt3D;
    o.vp = c.vp3D; o.vs = c.vs3D; o.density = c.density3D;
    o.layer_vp = c.layerVp3D; o.layer_vs = c.layerVs3D; o.layer_density = c.layerDensity3D;
    o.max_vp = c.maxVp3D; o.frequency = c.frequency3D;
    o.model_kind = int(c.modelKind3D);
    o.interface_depth = c.interfaceDepth3D; o.layer_dip = c.layerDip3D; o.layer_thickness = c.layerThickness3D;
    o.fault_throw = c.faultThrow3D; o.anomaly_radius = c.anomalyRadius3D;
    o.gx = c.gx3D; o.gy = c.gy3D; o.gz = c.gz3D; o.ox = c.ox3D; o.oy = c.oy3D; o.oz = c.oz3D;
    o.cells = c.cells3D;
    return o;
}

inline void fromC(const seis3d_ctx& o, SimulationContext3D& c) {
    c.nx3D = o.nx; c.ny3D = o.ny; c.nz3D = o.nz; c.pml3D = o.pml;
    c.freeSurface3D = o.free_surface != 0; c.reflecting3D = o.reflecting != 0;
    c.spacing3D = o.spacing; c.cfl3D = o.cfl; c.dt3D = o.dt;
    c.manualTimeStep3D = o.manual_dt != 0; c.requestedDtMs3D = o.requested_dt_ms; c.safeDt3D = o.safe_dt;
    c.vp3D = o.vp; c.vs3D = o.vs; c.density3D = o.density;
    c.layerVp3D = o.layer_vp; c.layerVs3D = o.layer_vs; c.layerDensity3D = o.layer_density;
    c.maxVp3D = o.max_vp; c.frequency3D = o.frequency;
    c.modelKind3D = ModelKind3D(o.model_kind);
    c.interfaceDepth3D = o.interface_depth; c.layerDip3D = o.layer_dip; c.layerThickness3D = o.layer_thickness;
    c.faultThrow3D = o.fault_throw; c.anomalyRadius3D = o.anomaly_radius;
    c.gx3D = o.gx; c.gy3D = o.gy; c.gz3D = o.gz; c.ox3D = o.ox; c.oy3D = o.oy; c.oz3D = o.oz;
    c.cells3D = std::size_t(o.cells);
}

inline seis3d_source toC(const Source3D& s) {
    seis3d_source o{};
    o.x = s.x3D; o.y = s.y3D; o.z = s.z3D; o.kind = int(s.kind3D); o.amplitude = s.amplitude3D;
    for (int i = 0; i < 3; ++i) o.direction[i] = s.direction3D[i];
    return o;
}

inline seis3d_survey toC(const Survey3D& q) {
    seis3d_survey o{};
    o.receiver_count_x = q.receiverCountX3D; o.receiver_count_y = q.receiverCountY3D;
    o.shot_count_x = q.shotCountX3D; o.shot_count_y = q.shotCountY3D;
    for (int a = 0; a < 3; ++a) { o.receiver_origin[a] = q.receiverOrigin3D[a]; o.shot_origin[a] = q.shotOrigin3D[a]; }
    for (int a = 0; a < 2; ++a) { o.receiver_spacing[a] = q.receiverSpacing3D[a]; o.shot_spacing[a] = q.shotSpacing3D[a]; }
    return o;
}

inline void fromC(const seis3d_survey& o, Survey3D& q) {
    q.receiverCountX3D = o.receiver_count_x; q.receiverCountY3D = o.receiver_count_y;
    q.shotCountX3D = o.shot_count_x; q.shotCountY3D = o.shot_count_y;
    for (int a = 0; a < 3; ++a) { q.receiverOrigin3D[a] = o.receiver_origin[a]; q.shotOrigin3D[a] = o.shot_origin[a]; }
    for (int a = 0; a < 2; ++a) { q.receiverSpacing3D[a] = o.receiver_spacing[a]; q.shotSpacing3D[a] = o.shot_spacing[a]; }
}

inline std::vector<seis3d_receiver> toC(const std::vector<Receiver3D>& r) {
    std::vector<seis3d_receiver> o(r.size());
    for (std::size_t i = 0; i < r.size(); ++i) o[i] = {r[i].x3D, r[i].y3D, r[i].z3D};
    return o;
}

inline seis3d_gather_info toC(const Gather3D& g) {
    seis3d_gather_info o{};
    o.spacing = g.context3D.spacing3D; o.frequency = g.context3D.frequency3D; o.source = toC(g.source3D);
    o.world_origin[0] = g.survey3D.worldOrigin3D[0]; o.world_origin[1] = g.survey3D.worldOrigin3D[1];
    o.cmp_bin[0] = g.survey3D.cmpBinSize3D[0]; o.cmp_bin[1] = g.survey3D.cmpBinSize3D[1];
    o.shot_id = g.shotId3D; o.stride = g.stride3D; o.samples = g.sampleCount3D;
    o.sample_interval_us = g.sampleIntervalUs3D; o.coordinate_scalar = g.coordinateScalar3D;
    o.first_sample_time = g.firstSampleTime3D;
    return o;
}

inline seis3d_t* session(const GPUSimData3D& g) {
    if (!g.session3D) throw std::logic_error("Seis3D: simulation not initialized.");
    return g.session3D;
}

}  // namespace detail

// validateContext3D (cuda/Cuda_Seis3D.cu:433-483).
inline void validateContext3D(SimulationContext3D& c) {
    seis3d_ctx o = detail::toC(c);
    detail::check(seis3d_validate_ctx(&o));
    detail::fromC(o, c);
}

// estimateMemory3D (cuda/Cuda_Seis3D.cu:504-516).
inline MemoryBudget3D estimateMemory3D(const SimulationContext3D& c) {
    const seis3d_ctx o = detail::toC(c);
    seis3d_budget b{};
    detail::check(seis3d_estimate_memory(&o, &b));
    MemoryBudget3D m;
    m.fieldBytes3D = b.field_bytes; m.cpmlBytes3D = b.cpml_bytes; m.viewBytes3D = b.view_bytes;
    m.totalBytes3D = b.total_bytes;
    return m;
}

// initializeGpu3D (cuda/Cuda_Seis3D.cu:519-555).
inline void initializeGpu3D(GPUSimData3D& g, const SimulationContext3D& c) {
    if (g.session3D) throw std::logic_error("Seis3D: release the previous session before initializing.");
    const seis3d_ctx o = detail::toC(c);
    seis3d_t* t = nullptr;
    detail::check(seis3d_create(&o, &t));
    g.session3D = t;
    g.receiverCount3D = 0;
    std::uint64_t bytes = 0;
    if (seis3d_device_bytes(&o, &bytes) == SEIS3D_OK) g.allocatedBytes3D = std::size_t(bytes);
}

// releaseGpu3D (cuda/Cuda_Seis3D.cu:558-569): idempotent.
inline void releaseGpu3D(GPUSimData3D& g) noexcept {
    seis3d_destroy(g.session3D);
    g = GPUSimData3D{};
}

// resetGpu3D (cuda/Cuda_Seis3D.cu:572-577).
inline void resetGpu3D(GPUSimData3D& g, const SimulationContext3D&) { detail::check(seis3d_reset(detail::session(g))); }

// refreshCpmlCoefficients3D (cuda/Cuda_Seis3D.cu:580-599).
inline void refreshCpmlCoefficients3D(GPUSimData3D& g, const SimulationContext3D& c) {
    const seis3d_ctx o = detail::toC(c);
    detail::check(seis3d_refresh_cpml(detail::session(g), &o));
}

// Learned CPML profile (no reference counterpart): SEIS3D_PROFILE_PARAMS shape parameters for
// this session's CPML tables, nullptr for the reference profile; kept by refreshCpmlCoefficients3D.
inline void setCpmlProfile3D(GPUSimData3D& g, const double* profile) {
    detail::check(seis3d_set_cpml_profile(detail::session(g), profile));
}

// runStep3D (cuda/Cuda_Seis3D.cu:602-624).
inline void runStep3D(GPUSimData3D& g, const SimulationContext3D&, const Source3D& s, std::uint64_t step) {
    const seis3d_source o = detail::toC(s);
    detail::check(seis3d_step(detail::session(g), &o, step));
}

// Several runStep3D calls enqueued at once (the loop of advanceSimulation3D,
// src/Seis3D_Sim.cpp:125-146), sampling into the device gather when recording.
inline void runSteps3D(GPUSimData3D& g, const Source3D& s, std::uint64_t first, std::int64_t count) {
    const seis3d_source o = detail::toC(s);
    detail::check(seis3d_run(detail::session(g), &o, first, count));
}

// setReceivers3D (cuda/Cuda_Seis3D.cu:649-668).
inline void setReceivers3D(GPUSimData3D& g, const SimulationContext3D&, const Receiver3D* r, int count) {
    static_assert(sizeof(Receiver3D) == sizeof(seis3d_receiver), "receiver layout");
    detail::check(seis3d_set_receivers(detail::session(g), reinterpret_cast<const seis3d_receiver*>(r), count));
    g.receiverCount3D = count;
}

// sampleReceivers3D (cuda/Cuda_Seis3D.cu:671-681).
inline const float* sampleReceivers3D(GPUSimData3D& g, const SimulationContext3D&) {
    const float* out = nullptr;
    detail::check(seis3d_sample_receivers(detail::session(g), &out));
    return out;
}

// probeMaterial3D (cuda/Cuda_Seis3D.cu:684-693).
inline void probeMaterial3D(GPUSimData3D& g, const SimulationContext3D&, const Receiver3D& p, float values[3]) {
    const seis3d_receiver r{p.x3D, p.y3D, p.z3D};
    detail::check(seis3d_probe(detail::session(g), &r, values));
}

// uploadModelPlane3D / finishModelUpload3D / downloadModelPlane3D (cuda/Cuda_Seis3D.cu:696-733).
inline void uploadModelPlane3D(GPUSimData3D& g, const SimulationContext3D&, int plane, const float* values) {
    detail::check(seis3d_upload_plane(detail::session(g), plane, values));
}
inline void finishModelUpload3D(GPUSimData3D& g, const SimulationContext3D&) {
    detail::check(seis3d_finish_upload(detail::session(g)));
}
inline void downloadModelPlane3D(GPUSimData3D& g, const SimulationContext3D&, int plane, float* values) {
    detail::check(seis3d_download_plane(detail::session(g), plane, values));
}

// Snapshot of one padded field (reference layout gx*gy*gz, x fastest); replaces reading
// g.field3D[f] with cudaMemcpy in the reference's tests.
inline std::vector<float> downloadField3D(GPUSimData3D& g, const SimulationContext3D& c, Field3D f) {
    std::vector<float> out(std::size_t(c.gx3D) * c.gy3D * c.gz3D);
    detail::check(seis3d_download_field(detail::session(g), int(f), out.data()));
    return out;
}

inline void uploadField3D(GPUSimData3D& g, Field3D f, const std::vector<float>& values) {
    detail::check(seis3d_upload_field(detail::session(g), int(f), values.data()));
}

// Wait for the session stream (the reference's cudaStreamSynchronize(g.stream3D)).
inline void synchronize3D(GPUSimData3D& g) { detail::check(seis3d_synchronize(detail::session(g))); }

// makeViewLayout3D (cuda/Cuda_Seis3D.cu:486-501).
inline ViewLayout3D makeViewLayout3D(const SimulationContext3D& c) {
    const seis3d_ctx cc = detail::toC(c);
    seis3d_view_layout o{};
    detail::check(seis3d_make_view_layout(&cc, &o));
    ViewLayout3D l;
    for (int a = 0; a < 3; ++a) {
        l.sliceOffset3D[a] = std::size_t(o.slice_offset[a]);
        l.sliceWidth3D[a] = o.slice_width[a];
        l.sliceHeight3D[a] = o.slice_height[a];
        l.domainSize3D[a] = o.domain_size[a];
        l.domainOrigin3D[a] = o.domain_origin[a];
    }
    l.volumeOffset3D = std::size_t(o.volume_offset);
    l.totalSamples3D = std::size_t(o.total_samples);
    l.volumeX3D = o.volume_x; l.volumeY3D = o.volume_y; l.volumeZ3D = o.volume_z;
    return l;
}

// extractViews3D (cuda/Cuda_Seis3D.cu:627-646).  `out` is host memory here (the reference
// writes a mapped GL buffer): 2 floats per sample, volume ? totalSamples3D : volumeOffset3D samples.
inline void extractViews3D(GPUSimData3D& g, const SimulationContext3D&, float* out, const ViewLayout3D& layout,
                           const int slices[3], int component, int materialComponent, bool volume, float& maximum,
                           float materialRange[2], bool includeBoundary, float* velocityMaximum = nullptr) {
    const std::size_t samples = volume ? layout.totalSamples3D : layout.volumeOffset3D;
    const int32_t s[3] = {slices[0], slices[1], slices[2]};
    seis3d_view_stats st{};
    detail::check(seis3d_extract_views(detail::session(g), out, samples * 2, s, component, materialComponent, volume,
                                       includeBoundary, &st));
    maximum = st.maximum;
    materialRange[0] = st.material_min;
    materialRange[1] = st.material_max;
    if (velocityMaximum) *velocityMaximum = st.velocity_maximum;
}

// ---- acquisition and files (src/Seis3D/Acquisition3D.cpp, ShotRecordIo3D.cpp, ModelIo3D.cpp) ----

// fitSurvey3D (src/Seis3D_Sim.cpp:42-60): fit the spreads into the model.
inline void fitSurvey3D(const SimulationContext3D& c, Survey3D& q) {
    const seis3d_ctx cc = detail::toC(c);
    seis3d_survey o = detail::toC(q);
    detail::check(seis3d_fit_survey(&cc, &o));
    detail::fromC(o, q);
}

// configureReceivers3D (src/Seis3D/Acquisition3D.cpp:21-37): receivers y-major, then x.
inline std::vector<Receiver3D> configureReceivers3D(const SimulationContext3D& c, const Survey3D& q) {
    const seis3d_ctx cc = detail::toC(c);
    const seis3d_survey o = detail::toC(q);
    std::vector<seis3d_receiver> r(maxReceivers3D);
    int count = maxReceivers3D;   // capacity in, receiver count out
    detail::check(seis3d_survey_receivers(&cc, &o, r.data(), &count));
    std::vector<Receiver3D> out(static_cast<std::size_t>(count));
    for (int i = 0; i < count; ++i) out[std::size_t(i)] = {r[std::size_t(i)].x, r[std::size_t(i)].y, r[std::size_t(i)].z};
    return out;
}

// startGather3D (src/Seis3D/Acquisition3D.cpp:188-230) without the UI state: validates the
// source, spread, sample count and stride, and freezes interval, first-sample time and XY scalar.
inline Gather3D startGather3D(const SimulationContext3D& c, const Source3D& s, const Survey3D& q,
                              const std::vector<Receiver3D>& receivers, int shotId, int targetSamples, int stride) {
    const seis3d_ctx cc = detail::toC(c);
    const seis3d_source cs = detail::toC(s);
    const auto cr = detail::toC(receivers);
    seis3d_gather_info info{};
    detail::check(seis3d_gather_info_init(&cc, &cs, cr.data(), int(cr.size()), shotId, stride, targetSamples,
                                          q.worldOrigin3D, q.cmpBinSize3D, &info));
    Gather3D g;
    g.context3D = c; g.source3D = s; g.survey3D = q; g.receivers3D = receivers;
    g.targetSamples3D = targetSamples; g.stride3D = stride; g.shotId3D = shotId;
    g.sampleIntervalUs3D = info.sample_interval_us; g.coordinateScalar3D = info.coordinate_scalar;
    g.firstSampleTime3D = info.first_sample_time;
    g.samples3D.reserve(std::size_t(targetSamples) * receivers.size() * 3);
    return g;
}

// exportGather3D (src/Seis3D/Acquisition3D.cpp:157-185): <prefix>_<shot:04>_vx/vy/vz.sgy (SEG-Y
// rev1) and _geometry.csv in `directory`; never overwrites.
inline void exportGather3D(const Gather3D& g, const std::string& directory, const std::string& prefix) {
    if (g.sampleCount3D < 1) throw std::runtime_error("No 3D receiver samples to export.");
    if (g.samples3D.size() != std::size_t(g.sampleCount3D) * g.receivers3D.size() * 3)
        throw std::invalid_argument("Seis3D: gather samples do not match the sample count.");
    const seis3d_gather_info info = detail::toC(g);
    const auto cr = detail::toC(g.receivers3D);
    detail::check(seis3d_export_gather(&info, cr.data(), int(cr.size()), g.samples3D.data(), directory.c_str(),
                                       prefix.c_str()));
}

// readShotSegy3D (src/Seis3D/ShotRecordIo3D.cpp:66-141): one fixed-length single-component shot.
// Receivers are the ranks of the trace coordinates along each axis.
inline ImportedShot3D readShotSegy3D(const std::string& path) {
    seis3d_segy_info info{};
    detail::check(seis3d_segy_inspect(path.c_str(), &info));
    const std::size_t n = std::size_t(info.traces), ns = std::size_t(info.samples);
    std::vector<float> traces(n * ns);
    std::vector<double> xyz(n * 3);
    ImportedShot3D shot;
    shot.firstTimes3D.resize(n);
    std::vector<int32_t> ids(n);
    detail::check(seis3d_segy_read(path.c_str(), traces.data(), xyz.data(), shot.firstTimes3D.data(), ids.data()));
    shot.traceIds3D.assign(ids.begin(), ids.end());
    shot.component3D = info.component;
    const char* names[3] = {"Vx", "Vy", "Vz"};
    shot.componentName3D = names[info.component];
    const char* units[] = {"file units", "Pa", "V", "mV", "A", "m", "m/s", "m/s^2", "N", "W"};
    shot.amplitudeUnit3D = info.measurement_unit >= 1 && info.measurement_unit <= 9 ? units[info.measurement_unit] : units[0];
    shot.cartesian3D = info.coordinate_units == 1;
    const std::size_t slash = path.find_last_of("/\\");
    shot.fileName3D = slash == std::string::npos ? path : path.substr(slash + 1);
    Gather3D& g = shot.gather3D;
    g.sampleCount3D = g.targetSamples3D = info.samples;
    g.sampleIntervalUs3D = info.sample_interval_us;
    g.shotId3D = info.shot_id;
    g.samples3D.assign(n * ns * 3, 0.0f);
    for (std::size_t r = 0; r < n; ++r)
        for (std::size_t k = 0; k < ns; ++k) g.samples3D[(k * n + r) * 3 + std::size_t(info.component)] = traces[r * ns + k];
    shot.coordinates3D.resize(n);
    g.receivers3D.resize(n);
    for (std::size_t r = 0; r < n; ++r) shot.coordinates3D[r] = {xyz[3 * r], xyz[3 * r + 1], xyz[3 * r + 2]};
    for (int axis = 0; axis < 3; ++axis) {
        std::vector<double> v(n);
        for (std::size_t r = 0; r < n; ++r) v[r] = shot.coordinates3D[r][std::size_t(axis)];
        std::vector<double> u = v;
        std::sort(u.begin(), u.end());
        u.erase(std::unique(u.begin(), u.end()), u.end());
        for (std::size_t r = 0; r < n; ++r) {
            const int rank = int(std::lower_bound(u.begin(), u.end(), v[r]) - u.begin());
            (axis == 0 ? g.receivers3D[r].x3D : axis == 1 ? g.receivers3D[r].y3D : g.receivers3D[r].z3D) = rank;
        }
    }
    g.firstSampleTime3D = shot.firstTimes3D.front();
    return shot;
}

// inspectModel3D (src/Seis3D/ModelIo3D.cpp:35-65): header, size and every cell of a .s3dm file.
inline SimulationContext3D inspectModel3D(const std::string& path, SimulationContext3D c) {
    seis3d_ctx cc = detail::toC(c);
    detail::check(seis3d_s3dm_inspect(path.c_str(), &cc));
    detail::fromC(cc, c);
    return c;
}

// importModel3D (:68-77): stream a validated model into a session of the same shape.
inline void importModel3D(GPUSimData3D& g, const SimulationContext3D&, const std::string& path) {
    detail::check(seis3d_s3dm_load(detail::session(g), path.c_str()));
}

// exportModel3D (:80-102): the session's model to a new file, written through <path>.part.
inline void exportModel3D(GPUSimData3D& g, const SimulationContext3D&, const std::string& path) {
    detail::check(seis3d_s3dm_save(detail::session(g), path.c_str()));
}

}  // namespace Seis3D
