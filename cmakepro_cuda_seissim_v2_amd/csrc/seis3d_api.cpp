// seis3d_api.cpp - C ABI and host session layer of the 3D solver.
//
// Restates the reference's Seis3D host API (cuda/Cuda_Seis3D.cu:433-733) and
// the cadence of its session driver (src/Seis3D_Sim.cpp:120-153,
// src/Seis3D/Acquisition3D.cpp:21-37,188-242) on HIP, with one non-blocking
// stream per session and no per-sample host synchronisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <tuple>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/seis3d.h"
#include "seis3d_host.hpp"
#include "seis3d_kernels.hpp"

namespace seis3d::host {

// validMaterial3D (cuda/Cuda_Seis3D.cu:30-34).
bool valid_material(float vp, float vs, float rho) {
    return std::isfinite(vp) && std::isfinite(vs) && std::isfinite(rho) && vp > 0 && vp <= 30000 && vs >= 0 &&
           vs < vp && vp * vp > (4.0f / 3.0f) * vs * vs && rho > 0 && rho <= 50000;
}


// validateContext3D (cuda/Cuda_Seis3D.cu:433-483).
void validate(seis3d_ctx& c) {
    if (c.reflecting) c.free_surface = 0;
    if (c.model_kind < 0 || c.model_kind > SEIS3D_MODEL_SALT_DOME || !std::isfinite(c.interface_depth) ||
        c.interface_depth < .1f || c.interface_depth > .9f || !std::isfinite(c.layer_dip) ||
        std::abs(c.layer_dip) > .6f || !std::isfinite(c.layer_thickness) || c.layer_thickness < .03f ||
        c.layer_thickness > .5f || !std::isfinite(c.fault_throw) || std::abs(c.fault_throw) > .4f ||
        !std::isfinite(c.anomaly_radius) || c.anomaly_radius < .05f || c.anomaly_radius > .4f)
        invalid("Seis3D: invalid procedural model geometry.");
    if (c.nx < 16 || c.ny < 16 || c.nz < 16) invalid("Seis3D: every physical dimension must be at least 16.");
    if (std::size_t(c.nx) > kMaxPhysicalNodes / std::size_t(c.ny) ||
        std::size_t(c.nx) * c.ny > kMaxPhysicalNodes / std::size_t(c.nz))
        invalid("Seis3D: physical node count exceeds 300 x 300 x 400 (36,000,000).");
    if (c.pml < 8 || c.pml > 40 || !std::isfinite(c.spacing) || c.spacing <= 0 || !std::isfinite(c.frequency) ||
        c.frequency <= 0 || c.frequency > 10000 || !std::isfinite(c.cfl) || c.cfl <= 0 || c.cfl > 0.9f)
        invalid("Seis3D: invalid CPML, spacing, frequency or CFL safety factor.");
    if (!valid_material(c.vp, c.vs, c.density) || !valid_material(c.layer_vp, c.layer_vs, c.layer_density))
        invalid("Seis3D: Vp/Vs/rho must be finite with positive bulk modulus.");
    if (c.model_kind != SEIS3D_MODEL_IMPORTED)
        c.max_vp = c.model_kind == SEIS3D_MODEL_UNIFORM ? c.vp : std::max(c.vp, c.layer_vp);
    if (!std::isfinite(c.max_vp) || c.max_vp <= 0 || c.max_vp > 30000) invalid("Seis3D: invalid model maximum Vp.");
    const int pml = active_pml(c);
    c.ox = c.oy = pml + 4;
    c.oz = c.free_surface ? 4 : pml + 4;
    c.gx = c.nx + 2 * pml + 8;
    c.gy = c.ny + 2 * pml + 8;
    c.gz = c.nz + (c.free_surface ? 1 : 2) * pml + 8;
    c.cells = std::uint64_t(c.gx) * c.gy * c.gz;
    constexpr double kSumWeights = 1225.0 / 1024 + 245.0 / 3072 + 49.0 / 5120 + 5.0 / 7168;
    const double safe = c.cfl * c.spacing / (std::sqrt(3.0) * c.max_vp * kSumWeights);
    const double us = std::floor(safe * 1.0e6);   // integer-microsecond dt for exact SEG-Y intervals
    if (us < 1 || us > 65535) invalid("Seis3D: safe dt is outside the implemented 1..65535 microsecond range.");
    c.safe_dt = float(us * 1.0e-6);
    c.dt = c.safe_dt;
    if (c.manual_dt) {
        const double requested = double(c.requested_dt_ms) * 1000;
        const double rounded = std::round(requested);
        if (!std::isfinite(requested) || rounded < 1 || rounded > 65535)
            invalid("Seis3D: manual dt must be between 0.001 and 65.535 ms.");
        if (std::abs(requested - rounded) > 0.05)
            invalid("Seis3D: manual dt must use whole microseconds (0.001 ms increments).");
        if (rounded > us)
            invalid("Seis3D: manual dt exceeds the model CFL ceiling; reduce dt or select Automatic CFL.");
        c.dt = float(rounded * 1.0e-6);
    }
}

// slabCount3D (cuda/Cuda_Seis3D.cu:37-42): reference slab cell counts.
std::size_t slab_count(const seis3d_ctx& c, int axis) {
    if (c.reflecting) return 0;
    if (axis == 0) return std::size_t(2 * c.pml) * c.gy * c.gz;
    if (axis == 1) return std::size_t(2 * c.pml) * c.gx * c.gz;
    return std::size_t((c.free_surface ? 1 : 2) * c.pml) * c.gx * c.gy;
}

// makeViewLayout3D (cuda/Cuda_Seis3D.cu:486-501): display domain sizes.
struct ViewLayout {
    std::size_t slice_offset[3]{};
    std::size_t volume_offset = 0, total = 0;
    int domain[3]{}, origin[3]{};
    int volume[3]{};
};
ViewLayout make_view_layout(const seis3d_ctx& c) {
    ViewLayout l;
    const int pml = active_pml(c);
    const int nx = c.nx + 2 * pml, ny = c.ny + 2 * pml, nz = c.nz + (c.free_surface ? 1 : 2) * pml;
    l.domain[0] = nx; l.domain[1] = ny; l.domain[2] = nz;
    l.origin[0] = -pml; l.origin[1] = -pml; l.origin[2] = c.free_surface ? 0 : -pml;
    l.slice_offset[1] = std::size_t(nx) * nz;
    l.slice_offset[2] = l.slice_offset[1] + std::size_t(ny) * nz;
    l.volume_offset = l.slice_offset[2] + std::size_t(nx) * ny;
    l.volume[0] = std::min(kMaxVolumeAxis, nx);
    l.volume[1] = std::min(kMaxVolumeAxis, ny);
    l.volume[2] = std::min(kMaxVolumeAxis, nz);
    l.total = l.volume_offset + std::size_t(l.volume[0]) * l.volume[1] * l.volume[2];
    return l;
}

// makeCoefficients3D (cuda/Cuda_Seis3D.cu:45-64); with a learned profile, the same layer depth,
// a/b/kappa formulas and table layout with the shapes of seis3d_set_cpml_profile.
std::vector<float> make_coefficients(const seis3d_ctx& c, int axis, const double* profile) {
    const int n = axis == 0 ? c.gx : axis == 1 ? c.gy : c.gz;
    std::vector<float> values(std::size_t(n) * 6);
    double sigma_max = -4.0 * c.max_vp * std::log(1.0e-6) / (2.0 * c.pml * c.spacing);
    if (profile) sigma_max *= std::exp(profile[0]);
    for (int half = 0; half < 2; ++half)
        for (int i = 0; i < n; ++i) {
            const double coordinate = i + 0.5 * half;
            double depth = (coordinate - (n - kRadius - c.pml - 0.5)) / c.pml;
            if (!(axis == 2 && c.free_surface)) depth = std::max(depth, (kRadius + c.pml - 0.5 - coordinate) / c.pml);
            depth = std::clamp(depth, 0.0, 1.0);
            double power, sigma, kappa, alpha;
            if (!profile) {
                power = depth * depth * depth;
                sigma = sigma_max * power;
                kappa = 1.0 + 2.0 * power;
                alpha = depth > 0 ? 3.141592653589793 * c.frequency * (1.0 - depth) : 0.0;
            } else {
                power = std::pow(depth, profile[1]);
                sigma = sigma_max * power;
                kappa = 1.0 + (profile[2] - 1.0) * power;
                alpha = depth > 0 ? profile[3] * 3.141592653589793 * c.frequency * std::pow(1.0 - depth, profile[4])
                                  : 0.0;
            }
            const double b = std::exp(-(sigma / kappa + alpha) * c.dt);
            const double a = sigma > 0 ? sigma * (b - 1.0) / (kappa * (sigma + kappa * alpha)) : 0.0;
            const std::size_t o = (std::size_t(half) * n + i) * 3;
            values[o] = float(a);
            values[o + 1] = float(b);
            values[o + 2] = float(1.0 / kappa);
        }
    return values;
}

void check_profile(const double* p) {
    if (!p) invalid("Seis3D: null CPML profile.");
    for (int k = 0; k < SEIS3D_PROFILE_PARAMS; ++k)
        if (!std::isfinite(p[k])) invalid("Seis3D: CPML profile parameters must be finite.");
    if (std::fabs(p[0]) > 5 || p[1] < 1 || p[1] > 8 || p[2] < 1 || p[2] > 20 || p[3] < 0 || p[3] > 10 ||
        p[4] < 0.1 || p[4] > 8)
        invalid("Seis3D: CPML profile parameters out of range.");
}

// Ricker sample of sourceKernel3D (cuda/Cuda_Seis3D.cu:319-320): FP64 phase,
// rounded to float.
float ricker(float frequency, double time) {
    const double phase = 3.141592653589793 * frequency * (time - 1.2 / frequency);
    return float((1 - 2 * phase * phase) * std::exp(-phase * phase));
}


Plan plan_for(const seis3d_ctx& c) {
    Plan p;
    p.px = int(align_up(std::size_t(c.gx), 64));
    p.plane = std::size_t(p.px) * c.gy;
    p.cells = p.plane * c.gz;
    const int pml = active_pml(c);
    if (pml > 0) {
        p.slab[0] = std::size_t(c.gz) * c.gy * (2 * pml);
        p.slab[1] = std::size_t(c.gz) * (2 * pml) * p.px;
        p.slab[2] = std::size_t((c.free_surface ? 1 : 2) * pml) * c.gy * p.px;
    }
    p.bytes = p.cells * 13 * sizeof(float);
    for (int a = 0; a < 3; ++a) p.bytes += p.slab[a] * 6 * sizeof(float);
    p.bytes += std::size_t(c.gx + c.gy + c.gz) * 6 * sizeof(float);
    return p;
}

}  // namespace seis3d::host


namespace seis3d::host {

void require_session(const seis3d_t* t) {
    if (!t || !t->stream || t->destroy_pending) fail(SEIS3D_ELIFECYCLE, "Seis3D: simulation not initialized.");
}

void require_unbatched(const seis3d_t* t, const char* operation) {
    require_session(t);
    if (t->batch_refs > 0)
        fail(SEIS3D_ELIFECYCLE, std::string("Seis3D: ") + operation +
                                    " is not allowed while a batch holds the session; destroy the batch first.");
}

void alloc_floats(float*& p, std::size_t count, hipStream_t s, const char* what) {
    if (count == 0) return;
    check(hipMalloc(&p, count * sizeof(float)), what);
    check(hipMemsetAsync(p, 0, count * sizeof(float), s), what);
}

void alloc_grid(float*& p, std::size_t cells, hipStream_t s, const char* what) {
    float* raw = nullptr;
    check(hipMalloc(&raw, (cells + 2 * kGridLead) * sizeof(float)), what);
    check(hipMemsetAsync(raw, 0, (cells + 2 * kGridLead) * sizeof(float), s), what);
    p = raw + kGridLead;
}

void upload_coefficients(seis3d_t* t, const seis3d_ctx& c, float* const dst[3]) {
    for (int a = 0; a < 3; ++a) {
        const auto v = make_coefficients(c, a, t->cpml_profile());
        check(hipMemcpyAsync(dst[a], v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice, t->stream),
              "upload CPML coefficients");
    }
    check(hipStreamSynchronize(t->stream), "finish coefficient upload");
}

void check_source(const seis3d_ctx& c, const seis3d_source& s) {
    if (s.x < 1 || s.x >= c.nx - 1 || s.y < 1 || s.y >= c.ny - 1 || s.z < 2 || s.z >= c.nz - 1 ||
        !std::isfinite(s.amplitude))
        invalid("Seis3D: source must be inside the physical volume, at depth >=2.");
    if (s.kind != SEIS3D_EXPLOSION && s.kind != SEIS3D_FORCE) invalid("Seis3D: invalid source kind.");
    for (float d : s.direction)
        if (!std::isfinite(d)) invalid("Seis3D: invalid force direction.");
}

// Source term at time `time` (sourceKernel3D, cuda/Cuda_Seis3D.cu:313-332).
seis3d::DevSource make_source(const seis3d_ctx& c, const seis3d_source& s, double time) {
    seis3d::DevSource d{};
    d.active = 1;
    d.x = s.x + c.ox; d.y = s.y + c.oy; d.z = s.z + c.oz;
    d.scale = c.dt * s.amplitude * ricker(c.frequency, time);
    for (int r = 0; r < 4; ++r) d.w[r] = std::exp(-0.5f * float(r));
    d.norm = kSourceNorm;
    for (int v = 0; v < 3; ++v) d.dir[v] = s.direction[v];
    return d;
}

void default_tiles(seis3d_t* t) {
    // LDS-DMA 2.5D family, 64 x 4 tiles, 16-plane chunks: the fastest shape
    // measured on MI355X for the 300 x 50 x 300 and 200^3 grids.
    t->kernels.family = 4;
    t->kernels.tile_y = 4;
    t->kernels.z_chunk = std::min(16, std::max(1, t->c.gz - 8));
}

// One runStep3D (cuda/Cuda_Seis3D.cu:602-624), enqueued on the session stream.
// Optional events bracket the two fused half-steps.
void enqueue_step(seis3d_t* t, const seis3d_source& s, std::uint64_t n, hipEvent_t* ev) {
    const seis3d_ctx& c = t->c;
    seis3d::DevStep P = t->base();
    if (t->inv_mu_dirty) {
        check(seis3d::launch_reciprocal_mu(P, t->stream), "kernel launch");
        t->inv_mu_dirty = false;
    }
    if (c.free_surface && !t->ghosts_valid) check(seis3d::launch_stress_ghost(P, t->stream), "kernel launch");
    if (c.reflecting) check(seis3d::launch_reflecting_walls(P, false, t->stream), "kernel launch");
    seis3d::DevStep V = P;
    if (s.kind == SEIS3D_FORCE) V.src = make_source(c, s, double(n) * c.dt);
    if (ev) check(hipEventRecord(ev[0], t->stream), "record event");
    check(seis3d::launch_velocity(V, t->kernels, t->stream), "kernel launch");
    if (ev) check(hipEventRecord(ev[1], t->stream), "record event");
    if (c.free_surface) check(seis3d::launch_velocity_ghost(P, c.spacing, t->stream), "kernel launch");
    if (c.reflecting) check(seis3d::launch_reflecting_walls(P, true, t->stream), "kernel launch");
    seis3d::DevStep S = P;
    if (s.kind == SEIS3D_EXPLOSION) S.src = make_source(c, s, (double(n) + 0.5) * c.dt);
    if (ev) check(hipEventRecord(ev[2], t->stream), "record event");
    check(seis3d::launch_stress(S, t->kernels, t->stream), "kernel launch");
    if (ev) check(hipEventRecord(ev[3], t->stream), "record event");
    if (t->abc) {
        enqueue_abc(t);
        if (c.free_surface) check(seis3d::launch_velocity_ghost(P, c.spacing, t->stream), "kernel launch");
    }
    // The family-4 stress kernel writes the stress ghost planes itself; the learned
    // frame changes stresses after it, so then the ghosts are redone here.
    if (c.free_surface && (t->kernels.family != 4 || t->abc))
        check(seis3d::launch_stress_ghost(P, t->stream), "kernel launch");
    if (c.reflecting) check(seis3d::launch_reflecting_walls(P, false, t->stream), "kernel launch");
    t->ghosts_valid = true;
}

// Sampling cadence of advanceSimulation3D (src/Seis3D_Sim.cpp:127-133).
void maybe_sample(seis3d_t* t, std::uint64_t n) {
    if (!t->recording || t->gather_count >= t->gather_target) return;
    if ((n + 1) % std::uint64_t(t->gather_stride) != 0) return;
    float* dst = t->gather + std::size_t(t->gather_count) * t->gather_receivers * 3;
    check(seis3d::launch_receivers(t->base(), t->receivers, t->receiver_count, dst, t->stream), "kernel launch");
    ++t->gather_count;
    if (t->gather_count >= t->gather_target) t->recording = false;
}

void read_flag(seis3d_t* t, unsigned int* flag) {
    check(hipMemcpyAsync(flag, t->flags, sizeof(unsigned int), hipMemcpyDeviceToHost, t->stream),
          "check capture stability");
}

long long view_samples(const seis3d_ctx& c, int volume) {
    const auto l = make_view_layout(c);
    return static_cast<long long>(volume ? l.total : l.volume_offset);
}

// extractViews3D (cuda/Cuda_Seis3D.cu:627-646) through device staging the caller provides:
// `stage` holds 2 floats per sample, `stat_words` 5 words.  Synchronises the session stream.
void extract_views_staged(seis3d_t* t, float* out, size_t capacity, const int32_t slices[3], int component,
                          int material_component, int volume, int include_boundary, seis3d_view_stats* stats,
                          float* stage, unsigned* stat_words) {
    const seis3d_ctx& c = t->c;
    if (!out || !slices || material_component < 5 || material_component > 7 || component < 0 || component > 7 ||
        slices[0] < 0 || slices[0] >= c.nx || slices[1] < 0 || slices[1] >= c.ny || slices[2] < 0 ||
        slices[2] >= c.nz)
        invalid("Seis3D: invalid view parameters.");
    const auto l = make_view_layout(c);
    const long long count = view_samples(c, volume);
    if (capacity < std::size_t(count) * 2) invalid("Seis3D: view buffer too small.");
    seis3d::ViewParams vp{};
    for (int a = 0; a < 3; ++a) {
        vp.slice_offset[a] = static_cast<long long>(l.slice_offset[a]);
        vp.volume[a] = l.volume[a];
        vp.domain[a] = l.domain[a];
        vp.origin[a] = l.origin[a];
        vp.slices[a] = slices[a];
    }
    vp.volume_offset = static_cast<long long>(l.volume_offset);
    vp.total = static_cast<long long>(l.total);
    vp.nx = c.nx; vp.ny = c.ny; vp.nz = c.nz;
    vp.ox = c.ox; vp.oy = c.oy; vp.oz = c.oz;
    vp.component = component;
    vp.material_component = material_component;
    vp.include_boundary = include_boundary != 0;
    static const unsigned init[5] = {0u, 0u, 0x7f7fffffu, 0u, 0u};
    check(hipMemcpyAsync(stat_words, init, sizeof(init), hipMemcpyHostToDevice, t->stream), "reset view statistics");
    check(seis3d::launch_extract_views(t->base(), vp, count, stage, stat_words, t->stream), "kernel launch");
    check(hipMemcpyAsync(out, stage, std::size_t(count) * 2 * sizeof(float), hipMemcpyDeviceToHost, t->stream),
          "read view");
    unsigned host[5]{};
    check(hipMemcpyAsync(host, stat_words, sizeof(host), hipMemcpyDeviceToHost, t->stream), "read diagnostics");
    check(hipStreamSynchronize(t->stream), "finish view extraction");
    if (host[1]) fail(SEIS3D_ENONFINITE, "Seis3D: non-finite wavefield detected; session stopped.");
    if (stats) {
        std::memcpy(&stats->maximum, host + 0, 4);
        std::memcpy(&stats->material_min, host + 2, 4);
        std::memcpy(&stats->material_max, host + 3, 4);
        std::memcpy(&stats->velocity_maximum, host + 4, 4);
    }
}

}  // namespace seis3d::host

using namespace seis3d::host;

extern "C" {

void seis3d_default_ctx(seis3d_ctx* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->nx = 300; c->ny = 50; c->nz = 300;
    c->pml = 12;
    c->free_surface = 1;
    c->reflecting = 0;
    c->spacing = 10.0f; c->cfl = 0.85f; c->dt = 0.0f;
    c->manual_dt = 0;
    c->requested_dt_ms = 0.5f;
    c->safe_dt = 0.0f;
    c->vp = 3000.0f; c->vs = 1700.0f; c->density = 2200.0f;
    c->layer_vp = 4200.0f; c->layer_vs = 2400.0f; c->layer_density = 2500.0f;
    c->max_vp = 4200.0f; c->frequency = 20.0f;
    c->model_kind = SEIS3D_MODEL_TWO_LAYER;
    c->interface_depth = 0.5f; c->layer_dip = 0.25f; c->layer_thickness = 0.15f;
    c->fault_throw = 0.18f; c->anomaly_radius = 0.187f;
}

void seis3d_default_source(seis3d_source* s) {
    if (!s) return;
    s->x = 48; s->y = 48; s->z = 24;
    s->kind = SEIS3D_EXPLOSION;
    s->amplitude = 1.0e8f;
    s->direction[0] = 0.0f; s->direction[1] = 0.0f; s->direction[2] = 1.0f;
}

void seis3d_default_survey(seis3d_survey* q) {
    if (!q) return;
    q->receiver_count_x = 50; q->receiver_count_y = 3;
    q->receiver_origin[0] = 27; q->receiver_origin[1] = 1; q->receiver_origin[2] = 0;
    q->receiver_spacing[0] = 5; q->receiver_spacing[1] = 23;
    q->shot_count_x = 3; q->shot_count_y = 2;
    q->shot_origin[0] = 24; q->shot_origin[1] = 24; q->shot_origin[2] = 12;
    q->shot_spacing[0] = 20; q->shot_spacing[1] = 20;
}

int seis3d_validate_ctx(seis3d_ctx* c) {
    return guarded([&] {
        if (!c) invalid("Seis3D: null context.");
        validate(*c);
    });
}

// estimateMemory3D (cuda/Cuda_Seis3D.cu:504-516): the reference's budget,
// including its GL staging, for callers that compare budgets.
int seis3d_estimate_memory(const seis3d_ctx* ctx, seis3d_budget* b) {
    return guarded([&] {
        if (!ctx || !b) invalid("Seis3D: null argument.");
        seis3d_ctx c = *ctx;
        validate(c);
        b->field_bytes = c.cells * 12 * sizeof(float);
        b->cpml_bytes = 0;
        for (int a = 0; a < 3; ++a) b->cpml_bytes += slab_count(c, a) * 6 * sizeof(float);
        const auto l = make_view_layout(c);
        b->view_bytes = l.total * 4 * sizeof(float) + l.volume_offset * 4 +
                        std::size_t(kMaxVolumeOutput) * kMaxVolumeOutput * 4;
        b->total_bytes = b->field_bytes + b->cpml_bytes + b->view_bytes +
                         std::size_t(c.gx + c.gy + c.gz) * 6 * sizeof(float) +
                         kMaxReceivers * (sizeof(seis3d_receiver) + 3 * sizeof(float)) + 2 * 1024 * 1024;
    });
}

int seis3d_device_bytes(const seis3d_ctx* ctx, uint64_t* out) {
    return guarded([&] {
        if (!ctx || !out) invalid("Seis3D: null argument.");
        seis3d_ctx c = *ctx;
        validate(c);
        *out = plan_for(c).bytes;
    });
}

int seis3d_cpml_coefficients(const seis3d_ctx* ctx, int axis, float* out, size_t capacity) {
    return guarded([&] {
        if (!ctx || !out || axis < 0 || axis > 2) invalid("Seis3D: invalid coefficient request.");
        const auto v = make_coefficients(*ctx, axis);
        if (capacity < v.size()) invalid("Seis3D: coefficient buffer too small.");
        std::copy(v.begin(), v.end(), out);
    });
}

// fitSurvey3D (src/Seis3D_Sim.cpp:42-60).
int seis3d_fit_survey(const seis3d_ctx* ctx, seis3d_survey* s) {
    return guarded([&] {
        if (!ctx || !s) invalid("Seis3D: null argument.");
        const seis3d_ctx& c = *ctx;
        s->receiver_count_x = std::clamp(s->receiver_count_x, 1, std::min(kMaxReceivers, c.nx - 2));
        s->receiver_count_y = std::clamp(s->receiver_count_y, 1, std::min(kMaxReceivers / s->receiver_count_x, c.ny - 2));
        s->shot_count_x = std::clamp(s->shot_count_x, 1, 100);
        s->shot_count_y = std::clamp(s->shot_count_y, 1, 100);
        s->receiver_origin[2] = 0;
        s->receiver_spacing[0] = std::max(1, (c.nx - 3) / std::max(1, s->receiver_count_x - 1));
        s->receiver_spacing[1] = std::max(1, (c.ny - 3) / std::max(1, s->receiver_count_y - 1));
        s->receiver_origin[0] = s->receiver_count_x == 1 ? c.nx / 2
                                : (c.nx - 1 - (s->receiver_count_x - 1) * s->receiver_spacing[0]) / 2;
        s->receiver_origin[1] = s->receiver_count_y == 1 ? c.ny / 2
                                : (c.ny - 1 - (s->receiver_count_y - 1) * s->receiver_spacing[1]) / 2;
        s->shot_origin[0] = c.nx / 4;
        s->shot_origin[1] = c.ny / 4;
        s->shot_origin[2] = std::max(2, c.nz / 8);
        s->shot_spacing[0] = std::max(1, (c.nx - 2 - s->shot_origin[0]) / std::max(1, s->shot_count_x - 1));
        s->shot_spacing[1] = std::max(1, (c.ny - 2 - s->shot_origin[1]) / std::max(1, s->shot_count_y - 1));
    });
}

// configureReceivers3D (src/Seis3D/Acquisition3D.cpp:21-37).
int seis3d_survey_receivers(const seis3d_ctx* ctx, const seis3d_survey* q, seis3d_receiver* out, int* count) {
    return guarded([&] {
        if (!ctx || !q || !count) invalid("Seis3D: null argument.");
        const seis3d_ctx& c = *ctx;
        if (q->receiver_count_x < 1 || q->receiver_count_y < 1 || q->receiver_count_x > kMaxReceivers ||
            q->receiver_count_y > kMaxReceivers ||
            std::int64_t(q->receiver_count_x) * q->receiver_count_y > kMaxReceivers || q->receiver_spacing[0] < 1 ||
            q->receiver_spacing[1] < 1)
            fail(SEIS3D_EHIP, "Receiver mesh must have 1..100 total channels (X times Y) and positive spacing.");
        const auto last_x = std::int64_t(q->receiver_origin[0]) + std::int64_t(q->receiver_count_x - 1) * q->receiver_spacing[0];
        const auto last_y = std::int64_t(q->receiver_origin[1]) + std::int64_t(q->receiver_count_y - 1) * q->receiver_spacing[1];
        if (q->receiver_origin[0] < 0 || q->receiver_origin[1] < 0 || last_x >= c.nx || last_y >= c.ny ||
            q->receiver_origin[2] < 0 || q->receiver_origin[2] >= c.nz)
            fail(SEIS3D_EHIP, "Receiver mesh exceeds the model.");
        const int n = q->receiver_count_x * q->receiver_count_y;
        if (!out || *count < n) { *count = n; invalid("Seis3D: receiver buffer too small."); }
        int k = 0;
        for (int y = 0; y < q->receiver_count_y; ++y)
            for (int x = 0; x < q->receiver_count_x; ++x)
                out[k++] = {q->receiver_origin[0] + x * q->receiver_spacing[0],
                            q->receiver_origin[1] + y * q->receiver_spacing[1], q->receiver_origin[2]};
        *count = n;
    });
}

// positionSequenceShot3D (src/Seis3D/Acquisition3D.cpp:233-242); src keeps
// its kind, amplitude and direction.
int seis3d_sequence_shot(const seis3d_ctx* ctx, const seis3d_survey* q, int index, seis3d_source* src) {
    return guarded([&] {
        if (!ctx || !q || !src) invalid("Seis3D: null argument.");
        const seis3d_ctx& c = *ctx;
        if (q->shot_count_x < 1 || q->shot_count_y < 1 || q->shot_count_x > 100 || q->shot_count_y > 100 ||
            q->shot_spacing[0] < 1 || q->shot_spacing[1] < 1)
            fail(SEIS3D_EHIP, "Invalid 3D source spread.");
        const auto last_x = std::int64_t(q->shot_origin[0]) + std::int64_t(q->shot_count_x - 1) * q->shot_spacing[0];
        const auto last_y = std::int64_t(q->shot_origin[1]) + std::int64_t(q->shot_count_y - 1) * q->shot_spacing[1];
        if (q->shot_origin[0] < 1 || q->shot_origin[1] < 1 || last_x >= c.nx - 1 || last_y >= c.ny - 1 ||
            q->shot_origin[2] < 2 || q->shot_origin[2] >= c.nz - 1)
            fail(SEIS3D_EHIP, "Source spread exceeds the 3D model.");
        if (index < 0 || index >= q->shot_count_x * q->shot_count_y) invalid("Seis3D: shot index out of range.");
        src->x = q->shot_origin[0] + (index % q->shot_count_x) * q->shot_spacing[0];
        src->y = q->shot_origin[1] + (index / q->shot_count_x) * q->shot_spacing[1];
        src->z = q->shot_origin[2];
    });
}

// startGather3D timing (src/Seis3D/Acquisition3D.cpp:201-209).
int seis3d_gather_timing(const seis3d_ctx* c, int stride, int* interval_us, double* first_sample_time) {
    return guarded([&] {
        if (!c || stride < 1 || stride > 10000) invalid("Seis3D: stride must be 1..10000.");
        const auto interval = std::llround(double(c->dt) * 1e6) * stride;
        if (interval > kMaxSegyIntervalUs) fail(SEIS3D_EHIP, "SEG-Y rev1 sample interval must not exceed 32767 us.");
        if (interval_us) *interval_us = int(interval);
        if (first_sample_time) *first_sample_time = (stride - 0.5) * c->dt;
    });
}

float seis3d_ricker(float frequency, double time) { return ricker(frequency, time); }

// initializeGpu3D (cuda/Cuda_Seis3D.cu:519-555).
int seis3d_create(const seis3d_ctx* ctx, seis3d_t** out) {
    return guarded([&] {
        if (!ctx || !out) invalid("Seis3D: null argument.");
        if (*out) fail(SEIS3D_ELIFECYCLE, "Seis3D: release the previous session before initializing.");
        seis3d_ctx c = *ctx;
        validate(c);
        if (c.gx != ctx->gx || c.gy != ctx->gy || c.gz != ctx->gz || c.dt != ctx->dt)
            invalid("Seis3D: context must pass seis3d_validate_ctx before initialization.");
        auto* t = new seis3d_session();
        t->c = c;
        t->plan = plan_for(c);
        try {
            check(hipGetDevice(&t->device), "query current device");
            std::size_t free_bytes = 0, total_bytes = 0;
            check(hipMemGetInfo(&free_bytes, &total_bytes), "query device memory");
            const std::size_t reserve = std::max(std::size_t(512) * 1024 * 1024, total_bytes / 20);
            if (free_bytes < reserve || t->plan.bytes > free_bytes - reserve)
                fail(SEIS3D_EHIP, "Seis3D: device memory budget exceeded; reduce grid size or CPML thickness.");
            check(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking), "create stream");
            for (auto& f : t->fields) alloc_grid(f, t->plan.cells, t->stream, "allocate field");
            alloc_grid(t->inv_rho, t->plan.cells, t->stream, "allocate field");
            alloc_grid(t->lambda, t->plan.cells, t->stream, "allocate field");
            alloc_grid(t->mu, t->plan.cells, t->stream, "allocate field");
            alloc_grid(t->inv_mu, t->plan.cells, t->stream, "allocate field");
            if (!c.reflecting) {
                for (int a = 0; a < 3; ++a) {
                    alloc_floats(t->psi[a], t->plan.slab[a] * 6, t->stream, "allocate CPML");
                    const int n = a == 0 ? c.gx : a == 1 ? c.gy : c.gz;
                    alloc_floats(t->coef[a], std::size_t(n) * 6, t->stream, "allocate CPML coefficients");
                }
                upload_coefficients(t, c, t->coef);
            }
            check(hipMalloc(&t->flags, 4 * sizeof(unsigned int)), "allocate statistics");
            check(hipMemsetAsync(t->flags, 0, 4 * sizeof(unsigned int), t->stream), "clear statistics");
            check(seis3d::launch_generate_materials(t->base(), t->material_params(), t->stream), "kernel launch");
            check(hipStreamSynchronize(t->stream), "initialize model");
            default_tiles(t);
        } catch (...) {
            t->release();
            delete t;
            throw;
        }
        *out = t;
    });
}

void seis3d_destroy(seis3d_t* t) {
    if (!t) return;
    if (t->batch_refs > 0) {
        // A batch graph still holds the session's buffers: free them with the last batch.
        t->destroy_pending = true;
        return;
    }
    t->release();
    delete t;
}

int seis3d_get_ctx(const seis3d_t* t, seis3d_ctx* out) {
    return guarded([&] {
        if (!t || !out) invalid("Seis3D: null argument.");
        *out = t->c;
    });
}

// resetGpu3D (cuda/Cuda_Seis3D.cu:572-577).
int seis3d_reset(seis3d_t* t) {
    return guarded([&] {
        require_session(t);
        for (auto p : t->fields) check(hipMemsetAsync(p, 0, t->plan.cells * sizeof(float), t->stream), "reset wavefield");
        for (int a = 0; a < 3; ++a)
            if (t->psi[a]) check(hipMemsetAsync(t->psi[a], 0, t->plan.slab[a] * 6 * sizeof(float), t->stream), "reset CPML");
        check(hipMemsetAsync(t->flags, 0, 4 * sizeof(unsigned int), t->stream), "reset diagnostics");
        t->ghosts_valid = true;
    });
}

// refreshCpmlCoefficients3D (cuda/Cuda_Seis3D.cu:580-599): new tables are
// built and uploaded into fresh buffers before the old ones are swapped out.
int seis3d_refresh_cpml(seis3d_t* t, const seis3d_ctx* ctx) {
    return guarded([&] {
        require_unbatched(t, "a CPML refresh");
        if (!ctx) invalid("Seis3D: null context.");
        seis3d_ctx c = *ctx;
        validate(c);
        if (c.gx != t->c.gx || c.gy != t->c.gy || c.gz != t->c.gz || c.reflecting != t->c.reflecting ||
            c.free_surface != t->c.free_surface || c.pml != t->c.pml || c.dt != t->c.dt)
            invalid("Seis3D: a CPML refresh may only change the source frequency.");
        if (c.reflecting) { t->c = c; return; }
        if (t->abc) fail(SEIS3D_ELIFECYCLE, "Seis3D: detach the learned boundary before refreshing the CPML.");
        float* replacement[3]{};
        try {
            for (int a = 0; a < 3; ++a) {
                const int n = a == 0 ? c.gx : a == 1 ? c.gy : c.gz;
                alloc_floats(replacement[a], std::size_t(n) * 6, t->stream, "refresh CPML coefficients");
            }
            upload_coefficients(t, c, replacement);
        } catch (...) {
            (void)hipStreamSynchronize(t->stream);
            for (auto p : replacement) (void)hipFree(p);
            throw;
        }
        for (int a = 0; a < 3; ++a) {
            std::swap(t->coef[a], replacement[a]);
            (void)hipFree(replacement[a]);
        }
        t->c = c;
    });
}

int seis3d_set_cpml_profile(seis3d_t* t, const double* profile) {
    return guarded([&] {
        require_unbatched(t, "a CPML profile change");
        if (profile) check_profile(profile);
        if (t->abc) fail(SEIS3D_ELIFECYCLE, "Seis3D: detach the learned boundary before changing the CPML profile.");
        const bool had = t->has_profile;
        double old[SEIS3D_PROFILE_PARAMS];
        std::copy(t->profile, t->profile + SEIS3D_PROFILE_PARAMS, old);
        t->has_profile = profile != nullptr;
        if (profile) std::copy(profile, profile + SEIS3D_PROFILE_PARAMS, t->profile);
        if (t->c.reflecting) return;
        try {
            upload_coefficients(t, t->c, t->coef);
        } catch (...) {
            t->has_profile = had;
            std::copy(old, old + SEIS3D_PROFILE_PARAMS, t->profile);
            throw;
        }
    });
}

int seis3d_get_cpml_profile(const seis3d_t* t, double* profile, int* learned) {
    return guarded([&] {
        require_session(t);
        static const double reference[SEIS3D_PROFILE_PARAMS] = {0.0, 3.0, 3.0, 1.0, 1.0};
        if (profile) std::copy_n(t->has_profile ? t->profile : reference, SEIS3D_PROFILE_PARAMS, profile);
        if (learned) *learned = t->has_profile ? 1 : 0;
    });
}

int seis3d_cpml_profile_coefficients(const seis3d_ctx* ctx, const double* profile, int axis, float* out,
                                     size_t capacity) {
    return guarded([&] {
        if (!ctx || !out || axis < 0 || axis > 2) invalid("Seis3D: invalid coefficient request.");
        if (profile) check_profile(profile);
        const auto v = make_coefficients(*ctx, axis, profile);
        if (capacity < v.size()) invalid("Seis3D: coefficient buffer too small.");
        std::copy(v.begin(), v.end(), out);
    });
}

int seis3d_read_cpml(seis3d_t* t, int axis, float* out, size_t capacity) {
    return guarded([&] {
        require_session(t);
        if (axis < 0 || axis > 2 || !out) invalid("Seis3D: invalid coefficient request.");
        if (!t->coef[axis]) invalid("Seis3D: no CPML in reflecting mode.");
        const int n = axis == 0 ? t->c.gx : axis == 1 ? t->c.gy : t->c.gz;
        if (capacity < std::size_t(n) * 6) invalid("Seis3D: coefficient buffer too small.");
        check(hipMemcpyAsync(out, t->coef[axis], std::size_t(n) * 6 * sizeof(float), hipMemcpyDeviceToHost, t->stream),
              "read CPML coefficients");
        check(hipStreamSynchronize(t->stream), "read CPML coefficients");
    });
}

// uploadModelPlane3D (cuda/Cuda_Seis3D.cu:696-711).
int seis3d_upload_plane(seis3d_t* t, int z, const float* values) {
    return guarded([&] {
        require_unbatched(t, "a model upload");
        const seis3d_ctx& c = t->c;
        if (z < 0 || z >= c.nz || !values) invalid("Seis3D: invalid model plane.");
        const std::size_t n = std::size_t(c.nx) * c.ny;
        std::vector<float> material(n * 3);
        for (std::size_t i = 0; i < n; ++i) {
            const float vp = values[3 * i], vs = values[3 * i + 1], rho = values[3 * i + 2];
            if (!valid_material(vp, vs, rho) || vp > c.max_vp * (1 + 1.0e-6f))
                invalid("Seis3D: invalid or changed imported material.");
            material[i] = 1 / rho;
            material[n + i] = rho * (vp * vp - 2 * vs * vs);
            material[2 * n + i] = rho * vs * vs;
        }
        const std::size_t offset = (std::size_t(z + c.oz) * c.gy + c.oy) * t->plan.px + c.ox;
        float* dst[3] = {t->inv_rho, t->lambda, t->mu};
        for (int a = 0; a < 3; ++a)
            check(hipMemcpy2DAsync(dst[a] + offset, std::size_t(t->plan.px) * 4, material.data() + a * n,
                                   std::size_t(c.nx) * 4, std::size_t(c.nx) * 4, c.ny, hipMemcpyHostToDevice, t->stream),
                  "upload material plane");
        check(hipStreamSynchronize(t->stream), "finish material plane");
        t->inv_mu_dirty = true;
    });
}

// finishModelUpload3D (cuda/Cuda_Seis3D.cu:714-717).
int seis3d_finish_upload(seis3d_t* t) {
    return guarded([&] {
        require_unbatched(t, "a model upload");
        check(seis3d::launch_extend_materials(t->base(), t->material_params(), t->stream), "kernel launch");
        check(hipStreamSynchronize(t->stream), "finish model import");
        t->inv_mu_dirty = true;
    });
}

// downloadModelPlane3D (cuda/Cuda_Seis3D.cu:720-733).
int seis3d_download_plane(seis3d_t* t, int z, float* values) {
    return guarded([&] {
        require_session(t);
        const seis3d_ctx& c = t->c;
        if (z < 0 || z >= c.nz || !values) invalid("Seis3D: invalid export plane.");
        const std::size_t n = std::size_t(c.nx) * c.ny;
        std::vector<float> material(n * 3);
        const std::size_t offset = (std::size_t(z + c.oz) * c.gy + c.oy) * t->plan.px + c.ox;
        const float* src[3] = {t->inv_rho, t->lambda, t->mu};
        for (int a = 0; a < 3; ++a)
            check(hipMemcpy2DAsync(material.data() + a * n, std::size_t(c.nx) * 4, src[a] + offset,
                                   std::size_t(t->plan.px) * 4, std::size_t(c.nx) * 4, c.ny, hipMemcpyDeviceToHost,
                                   t->stream),
                  "download material plane");
        check(hipStreamSynchronize(t->stream), "finish model plane export");
        for (std::size_t i = 0; i < n; ++i) {
            values[3 * i] = std::sqrt((material[n + i] + 2 * material[2 * n + i]) * material[i]);
            values[3 * i + 1] = std::sqrt(material[2 * n + i] * material[i]);
            values[3 * i + 2] = 1 / material[i];
        }
    });
}

// probeMaterial3D (cuda/Cuda_Seis3D.cu:684-693).
int seis3d_probe(seis3d_t* t, const seis3d_receiver* p, float values[3]) {
    return guarded([&] {
        require_session(t);
        const seis3d_ctx& c = t->c;
        if (!p || !values) invalid("Seis3D: null argument.");
        if (p->x < 0 || p->x >= c.nx || p->y < 0 || p->y >= c.ny || p->z < 0 || p->z >= c.nz) return;
        const std::size_t i = (std::size_t(p->z + c.oz) * c.gy + (p->y + c.oy)) * t->plan.px + (p->x + c.ox);
        float inverse = 0, mu = 0, lambda = 0;
        check(hipMemcpyAsync(&inverse, t->inv_rho + i, 4, hipMemcpyDeviceToHost, t->stream), "probe density");
        check(hipMemcpyAsync(&mu, t->mu + i, 4, hipMemcpyDeviceToHost, t->stream), "probe mu");
        check(hipMemcpyAsync(&lambda, t->lambda + i, 4, hipMemcpyDeviceToHost, t->stream), "probe lambda");
        check(hipStreamSynchronize(t->stream), "finish material probe");
        values[0] = std::sqrt((lambda + 2 * mu) * inverse);
        values[1] = std::sqrt(mu * inverse);
        values[2] = 1 / inverse;
    });
}

int seis3d_step(seis3d_t* t, const seis3d_source* src, uint64_t step) {
    return guarded([&] {
        require_session(t);
        if (!src) invalid("Seis3D: null source.");
        check_source(t->c, *src);
        enqueue_step(t, *src, step);
    });
}

int seis3d_run(seis3d_t* t, const seis3d_source* src, uint64_t first_step, int64_t n_steps) {
    return guarded([&] {
        require_session(t);
        if (!src || n_steps < 0) invalid("Seis3D: invalid run request.");
        check_source(t->c, *src);
        for (int64_t k = 0; k < n_steps; ++k) {
            const std::uint64_t n = first_step + std::uint64_t(k);
            enqueue_step(t, *src, n);
            maybe_sample(t, n);
        }
    });
}

int seis3d_synchronize(seis3d_t* t) {
    return guarded([&] {
        require_session(t);
        unsigned int flag = 0;
        read_flag(t, &flag);
        check(hipStreamSynchronize(t->stream), "synchronize");
        if (flag) fail(SEIS3D_ENONFINITE, "Seis3D: non-finite wavefield detected; session stopped.");
    });
}

// setReceivers3D (cuda/Cuda_Seis3D.cu:649-668): validated, all-or-nothing.
int seis3d_set_receivers(seis3d_t* t, const seis3d_receiver* r, int count) {
    return guarded([&] {
        require_unbatched(t, "replacing receivers");
        const seis3d_ctx& c = t->c;
        if (count < 0 || count > kMaxReceivers || (count && !r)) invalid("Seis3D: invalid receiver count.");
        if (t->recording) fail(SEIS3D_ELIFECYCLE, "Seis3D: stop the gather before replacing receivers.");
        std::vector<int> padded(std::size_t(count) * 3);
        for (int a = 0; a < count; ++a) {
            const auto p = r[a];
            if (p.x < 0 || p.x >= c.nx || p.y < 0 || p.y >= c.ny || p.z < 0 || p.z >= c.nz)
                invalid("Seis3D: receiver outside the physical domain.");
            padded[3 * a] = p.x + c.ox;
            padded[3 * a + 1] = p.y + c.oy;
            padded[3 * a + 2] = p.z + c.oz;
        }
        int* positions = nullptr;
        float* values = nullptr;
        float* host = nullptr;
        try {
            if (count) {
                check(hipMalloc(&positions, padded.size() * sizeof(int)), "allocate receivers");
                check(hipMalloc(&values, std::size_t(count) * 3 * sizeof(float)), "allocate receiver values");
                check(hipHostMalloc(&host, std::size_t(count) * 3 * sizeof(float), hipHostMallocDefault),
                      "allocate pinned receivers");
                check(hipMemcpyAsync(positions, padded.data(), padded.size() * sizeof(int), hipMemcpyHostToDevice,
                                     t->stream),
                      "upload receivers");
            }
            check(hipStreamSynchronize(t->stream), "finish receiver update");
        } catch (...) {
            (void)hipFree(positions);
            (void)hipFree(values);
            if (host) (void)hipHostFree(host);
            throw;
        }
        (void)hipFree(t->receivers);
        (void)hipFree(t->receiver_values);
        if (t->host_values) (void)hipHostFree(t->host_values);
        t->receivers = positions;
        t->receiver_values = values;
        t->host_values = host;
        t->receiver_count = count;
    });
}

int seis3d_receiver_count(const seis3d_t* t) { return t ? t->receiver_count : 0; }

// sampleReceivers3D (cuda/Cuda_Seis3D.cu:671-681).
int seis3d_sample_receivers(seis3d_t* t, const float** out) {
    return guarded([&] {
        require_session(t);
        if (!out) invalid("Seis3D: null argument.");
        *out = nullptr;
        if (!t->receiver_count) return;
        check(seis3d::launch_receivers(t->base(), t->receivers, t->receiver_count, t->receiver_values, t->stream),
              "kernel launch");
        check(hipMemcpyAsync(t->host_values, t->receiver_values, std::size_t(t->receiver_count) * 3 * sizeof(float),
                             hipMemcpyDeviceToHost, t->stream),
              "gather receivers");
        unsigned int invalid_flag = 0;
        read_flag(t, &invalid_flag);
        check(hipStreamSynchronize(t->stream), "finish receiver gather");
        if (invalid_flag) fail(SEIS3D_ENONFINITE, "Seis3D: non-finite wavefield detected; capture stopped before export.");
        *out = t->host_values;
    });
}

// Recording state of startGather3D (src/Seis3D/Acquisition3D.cpp:188-230).
int seis3d_gather_begin(seis3d_t* t, int target_samples, int stride) {
    return guarded([&] {
        require_unbatched(t, "starting a gather");
        if (target_samples < 1 || target_samples > kMaxSegySamples || stride < 1 || stride > 10000)
            invalid("Gather sample count must be 1..32767 with stride 1..10000.");
        if (!t->receiver_count) invalid("Seis3D: set receivers before starting a gather.");
        const std::size_t values = std::size_t(target_samples) * t->receiver_count * 3;
        float* g = nullptr;
        check(hipMalloc(&g, values * sizeof(float)), "allocate gather");
        (void)hipFree(t->gather);
        t->gather = g;
        t->gather_target = target_samples;
        t->gather_stride = stride;
        t->gather_receivers = t->receiver_count;
        t->gather_count = 0;
        t->recording = true;
    });
}

int seis3d_gather_count(const seis3d_t* t) { return t ? t->gather_count : 0; }

int seis3d_gather_read(seis3d_t* t, float* out, size_t capacity, int* samples) {
    return guarded([&] {
        require_session(t);
        if (!t->gather) invalid("Seis3D: no gather recorded.");
        const std::size_t values = std::size_t(t->gather_count) * t->gather_receivers * 3;
        if (values && (!out || capacity < values)) invalid("Seis3D: gather buffer too small.");
        unsigned int flag = 0;
        read_flag(t, &flag);
        if (values)
            check(hipMemcpyAsync(out, t->gather, values * sizeof(float), hipMemcpyDeviceToHost, t->stream), "read gather");
        check(hipStreamSynchronize(t->stream), "finish gather read");
        if (flag) fail(SEIS3D_ENONFINITE, "Seis3D: non-finite wavefield detected; capture stopped before export.");
        if (samples) *samples = t->gather_count;
    });
}

int seis3d_gather_end(seis3d_t* t) {
    return guarded([&] {
        require_unbatched(t, "ending a gather");
        check(hipStreamSynchronize(t->stream), "finish gather");
        (void)hipFree(t->gather);
        t->gather = nullptr;
        t->recording = false;
        t->gather_count = t->gather_target = 0;
    });
}

namespace {
// The session's packed transfer buffer (one unpadded-pitch field), allocated on first use.
void ensure_stage(seis3d_t* t) {
    if (t->stage) return;
    check(hipMalloc(&t->stage, std::size_t(t->c.gx) * t->c.gy * t->c.gz * sizeof(float)), "allocate transfer buffer");
}
}  // namespace

int seis3d_download_field(seis3d_t* t, int field, float* out) {
    return guarded([&] {
        require_session(t);
        if (field < 0 || field > 11 || !out) invalid("Seis3D: invalid field request.");
        const float* src = field < 9 ? t->fields[field] : field == 9 ? t->inv_rho : field == 10 ? t->lambda : t->mu;
        const seis3d_ctx& c = t->c;
        // Pitched rows are packed on the device, then cross PCIe in one contiguous copy (a pitched
        // copy straight into pageable host memory moves row by row).
        const std::size_t rows = std::size_t(c.gy) * c.gz;
        ensure_stage(t);
        check(hipMemcpy2DAsync(t->stage, std::size_t(c.gx) * 4, src, std::size_t(t->plan.px) * 4,
                               std::size_t(c.gx) * 4, rows, hipMemcpyDeviceToDevice, t->stream),
              "download field");
        check(hipMemcpyAsync(out, t->stage, rows * c.gx * 4, hipMemcpyDeviceToHost, t->stream), "download field");
        check(hipStreamSynchronize(t->stream), "finish field download");
    });
}

int seis3d_upload_field(seis3d_t* t, int field, const float* in) {
    return guarded([&] {
        require_session(t);
        if (field < 0 || field > 8 || !in) invalid("Seis3D: invalid field request.");
        const seis3d_ctx& c = t->c;
        const std::size_t rows = std::size_t(c.gy) * c.gz;
        ensure_stage(t);
        check(hipMemcpyAsync(t->stage, in, rows * c.gx * 4, hipMemcpyHostToDevice, t->stream), "upload field");
        check(hipMemcpy2DAsync(t->fields[field], std::size_t(t->plan.px) * 4, t->stage, std::size_t(c.gx) * 4,
                               std::size_t(c.gx) * 4, rows, hipMemcpyDeviceToDevice, t->stream),
              "upload field");
        check(hipStreamSynchronize(t->stream), "finish field upload");
        t->ghosts_valid = false;
    });
}

int seis3d_extract_slice(seis3d_t* t, int axis, int at, int component, float* out, size_t capacity) {
    return guarded([&] {
        require_session(t);
        const seis3d_ctx& c = t->c;
        const int n[3] = {c.nx, c.ny, c.nz};
        if (axis < 0 || axis > 2 || component < 0 || component > 7 || !out || at < 0 || at >= n[axis])
            invalid("Seis3D: invalid view parameters.");
        const auto l = make_view_layout(c);
        seis3d::SliceParams sp{};
        sp.fixed_axis = axis;
        sp.fixed = at;
        sp.u_axis = axis == 0 ? 1 : 0;
        sp.v_axis = axis == 2 ? 1 : 2;
        sp.u_origin = l.origin[sp.u_axis];
        sp.v_origin = l.origin[sp.v_axis];
        sp.width = l.domain[sp.u_axis];
        sp.height = l.domain[sp.v_axis];
        sp.ox = c.ox; sp.oy = c.oy; sp.oz = c.oz;
        sp.component = component;
        const std::size_t count = std::size_t(sp.width) * sp.height;
        if (capacity < count) invalid("Seis3D: slice buffer too small.");
        float* d = nullptr;
        check(hipMalloc(&d, count * sizeof(float)), "allocate slice");
        try {
            check(seis3d::launch_extract_slice(t->base(), sp, d, t->stream), "kernel launch");
            check(hipMemcpyAsync(out, d, count * sizeof(float), hipMemcpyDeviceToHost, t->stream), "read slice");
            unsigned int flag = 0;
            read_flag(t, &flag);
            check(hipStreamSynchronize(t->stream), "finish view extraction");
            (void)hipFree(d);
            d = nullptr;
            if (flag) fail(SEIS3D_ENONFINITE, "Seis3D: non-finite wavefield detected; session stopped.");
        } catch (...) {
            (void)hipFree(d);
            throw;
        }
    });
}

int seis3d_make_view_layout(const seis3d_ctx* ctx, seis3d_view_layout* out) {
    return guarded([&] {
        if (!ctx || !out) invalid("Seis3D: null argument.");
        const auto l = make_view_layout(*ctx);
        seis3d_view_layout o{};
        for (int a = 0; a < 3; ++a) {
            o.slice_offset[a] = l.slice_offset[a];
            o.domain_size[a] = l.domain[a];
            o.domain_origin[a] = l.origin[a];
        }
        o.slice_width[0] = l.domain[0]; o.slice_height[0] = l.domain[2];
        o.slice_width[1] = l.domain[1]; o.slice_height[1] = l.domain[2];
        o.slice_width[2] = l.domain[0]; o.slice_height[2] = l.domain[1];
        o.volume_offset = l.volume_offset;
        o.total_samples = l.total;
        o.volume_x = l.volume[0]; o.volume_y = l.volume[1]; o.volume_z = l.volume[2];
        *out = o;
    });
}

int seis3d_extract_views(seis3d_t* t, float* out, size_t capacity, const int32_t slices[3], int component,
                         int material_component, int volume, int include_boundary, seis3d_view_stats* stats) {
    return guarded([&] {
        require_session(t);
        const long long count = view_samples(t->c, volume);
        float* d = nullptr;
        unsigned* st = nullptr;
        check(hipMalloc(&d, std::size_t(count) * 2 * sizeof(float)), "allocate view");
        try {
            check(hipMalloc(&st, 5 * sizeof(unsigned)), "allocate view statistics");
            extract_views_staged(t, out, capacity, slices, component, material_component, volume, include_boundary,
                                 stats, d, st);
        } catch (...) {
            (void)hipFree(d);
            (void)hipFree(st);
            throw;
        }
        (void)hipFree(d);
        (void)hipFree(st);
    });
}

int seis3d_set_tiles(seis3d_t* t, int family, int tile_y, int z_chunk) {
    return guarded([&] {
        require_session(t);
        if (family != 0 && family != 4) invalid("Seis3D: kernel family must be 0 (direct) or 4 (LDS-DMA).");
        if (family == 4 && tile_y != 4 && tile_y != 8) invalid("Seis3D: tile_y must be 4 or 8.");
        if (family == 4 && (z_chunk < 0 || z_chunk > 128 || (z_chunk == 0 && tile_y != 8)))
            invalid("Seis3D: z_chunk must be 1..128 for the LDS-DMA family (0: balanced one-round chunks, tile_y 8).");
        t->kernels.family = family;
        if (family >= 1) {
            t->kernels.tile_y = tile_y;
            t->kernels.z_chunk = z_chunk;
        }
    });
}

int seis3d_get_tiles(const seis3d_t* t, int* family, int* tile_y, int* z_chunk) {
    return guarded([&] {
        if (!t) invalid("Seis3D: null session.");
        if (family) *family = t->kernels.family;
        if (tile_y) *tile_y = t->kernels.tile_y;
        if (z_chunk) *z_chunk = t->kernels.z_chunk;
    });
}

int seis3d_run_timed(seis3d_t* t, const seis3d_source* src, uint64_t first_step, int64_t n_steps, seis3d_timing* out) {
    return guarded([&] {
        require_session(t);
        if (!src || n_steps < 1 || n_steps > 100000 || !out) invalid("Seis3D: invalid timed run request.");
        check_source(t->c, *src);
        std::vector<hipEvent_t> ev(std::size_t(n_steps) * 4 + 2, nullptr);
        try {
            for (auto& e : ev) check(hipEventCreate(&e), "create event");
            check(hipEventRecord(ev[0], t->stream), "record event");
            for (int64_t k = 0; k < n_steps; ++k) {
                const std::uint64_t n = first_step + std::uint64_t(k);
                enqueue_step(t, *src, n, ev.data() + 2 + 4 * k);
                maybe_sample(t, n);
            }
            check(hipEventRecord(ev[1], t->stream), "record event");
            check(hipEventSynchronize(ev[1]), "finish timed run");
            double v = 0, s = 0;
            float ms = 0;
            for (int64_t k = 0; k < n_steps; ++k) {
                hipEvent_t* e = ev.data() + 2 + 4 * k;
                check(hipEventElapsedTime(&ms, e[0], e[1]), "read event");
                v += ms;
                check(hipEventElapsedTime(&ms, e[2], e[3]), "read event");
                s += ms;
            }
            check(hipEventElapsedTime(&ms, ev[0], ev[1]), "read event");
            out->velocity_ms = v;
            out->stress_ms = s;
            out->total_ms = ms;
            out->launches = n_steps;
        } catch (...) {
            for (auto e : ev) if (e) (void)hipEventDestroy(e);
            throw;
        }
        for (auto e : ev) (void)hipEventDestroy(e);
    });
}

// Algorithmic bytes of one step: 30 floats per updated cell (velocity: 6
// stresses + 3 velocities + 1/rho read, 3 velocities written; stress: 3
// velocities + 6 stresses + lambda + mu read, 6 stresses written) and 12
// floats per (cell, axis) inside that axis' CPML slab (6 psi slots read and
// written, 3 in each half-step).
int seis3d_step_bytes(const seis3d_ctx* ctx, uint64_t* total, uint64_t* velocity, uint64_t* stress, uint64_t* n_upd,
                      uint64_t* n_psi) {
    return guarded([&] {
        if (!ctx) invalid("Seis3D: null context.");
        seis3d_ctx c = *ctx;
        validate(c);
        const std::uint64_t upd = std::uint64_t(c.gx - 8) * (c.gy - 8) * (c.gz - 8);
        std::uint64_t psi_cells = 0;
        const int pml = active_pml(c);
        if (pml) {
            psi_cells += std::uint64_t(2 * pml) * (c.gy - 8) * (c.gz - 8);
            psi_cells += std::uint64_t(2 * pml) * (c.gx - 8) * (c.gz - 8);
            psi_cells += std::uint64_t((c.free_surface ? 1 : 2) * pml) * (c.gx - 8) * (c.gy - 8);
        }
        const std::uint64_t v = 4 * (13 * upd + 6 * psi_cells);
        const std::uint64_t s = 4 * (17 * upd + 6 * psi_cells);
        if (total) *total = v + s;
        if (velocity) *velocity = v;
        if (stress) *stress = s;
        if (n_upd) *n_upd = upd;
        if (n_psi) *n_psi = psi_cells;
    });
}

const char* seis3d_last_error(void) { return g_error.c_str(); }

const char* seis3d_version(void) { return "seiswave-mi355x 0.1.0 (gfx950, HIP)"; }

int seis3d_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Autotuner and multi-shot batches
// ---------------------------------------------------------------------------

namespace {

using TileKey = std::tuple<int, int, int, int, int, int>;
struct Tuned {
    seis3d::KernelConfig config;
    double ms_per_step;
};
std::mutex g_tune_mutex;
std::map<TileKey, Tuned> g_tuned;

TileKey tile_key(const seis3d_ctx& c) { return {c.gx, c.gy, c.gz, c.pml, c.free_surface, c.reflecting}; }

// Candidate shapes: the direct family and the LDS-DMA family over tile
// heights and z-chunk lengths that keep at least a few hundred workgroups.
std::vector<seis3d::KernelConfig> tile_candidates(const seis3d_ctx& c) {
    std::vector<seis3d::KernelConfig> out;
    out.push_back({0, 8, 32});
    const int uz = c.gz - 8;
    for (int ty : {4, 8}) {
        std::vector<int> chunks = {8, 12, 16, 24, 32};
        // Chunk sizes whose launch fills whole rounds of resident workgroups
        // (the last round of a launch that overflows a round runs mostly idle).
        int per_cu = 0, cus = 0;
        if (seis3d::dma_residency(ty, &per_cu, &cus) == hipSuccess && per_cu > 0 && cus > 0) {
            const int layer = ((c.gx - 8 + 63) / 64) * ((c.gy - 8 + ty - 1) / ty);
            for (int rounds = 1; rounds <= 6; ++rounds) {
                const int nbz = rounds * per_cu * cus / layer;
                if (nbz >= 1) chunks.push_back((uz + nbz - 1) / nbz);
            }
        }
        std::sort(chunks.begin(), chunks.end());
        chunks.erase(std::unique(chunks.begin(), chunks.end()), chunks.end());
        for (int zc : chunks)
            if (zc >= 8 && zc <= uz && zc <= 128) out.push_back({4, ty, zc});
    }
    out.push_back({4, 8, 0});   // one round, chunk counts balanced across the tile columns
    return out;
}

}  // namespace

struct seis3d_batch {
    // Gather bookkeeping of one session: before the batch's steps and after them.  The
    // graph writes the same gather slots on every launch, so a launch sets `after`.
    struct Recording {
        int count = 0;
        bool recording = false;
        bool ghosts = true;
    };
    std::vector<seis3d_t*> sessions;
    std::vector<Recording> before, after;
    hipStream_t origin = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<hipEvent_t> events;
    int64_t steps = 0;
    bool holds = false;   // sessions' batch_refs were raised by this batch
    bool multi = false;   // one launch per half-step for every shot (else a stream per shot)
    // Device tables of the multi-shot ghost and receiver launches: per step and half-step every
    // shot's step parameters, every shot's spacing, and the receiver samples.  (The half-steps
    // themselves take their shots' parameters as kernel arguments.)
    seis3d::DevStep* d_steps = nullptr;
    float* d_spacing = nullptr;
    seis3d::ShotSample* d_samples = nullptr;
    // Work-queue mode: every shot's tile map, the shots' first item numbers, and eight queue
    // counters per launch (zeroed by the graph's first node).
    seis3d::TileMap* d_maps = nullptr;
    int* d_first = nullptr;
    unsigned* d_counters = nullptr;

    void release() noexcept {
        if (origin) (void)hipStreamSynchronize(origin);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto e : events) (void)hipEventDestroy(e);
        if (origin) (void)hipStreamDestroy(origin);
        exec = nullptr; graph = nullptr; origin = nullptr;
        events.clear();
        (void)hipFree(d_steps); (void)hipFree(d_spacing); (void)hipFree(d_samples);
        (void)hipFree(d_maps); (void)hipFree(d_first); (void)hipFree(d_counters);
        d_steps = nullptr; d_spacing = nullptr; d_samples = nullptr;
        d_maps = nullptr; d_first = nullptr; d_counters = nullptr;
        if (holds) {
            for (auto* t : sessions) {
                if (--t->batch_refs == 0 && t->destroy_pending) {
                    t->release();
                    delete t;
                }
            }
            holds = false;
        }
        sessions.clear();
    }
};

extern "C" {

int seis3d_autotune(seis3d_t* t, const seis3d_source* src, int steps, int* family, int* tile_y, int* z_chunk,
                    double* best_ms) {
    return guarded([&] {
        require_session(t);
        if (!src || steps < 1 || steps > 1000) invalid("Seis3D: invalid autotune request.");
        check_source(t->c, *src);
        const TileKey key = tile_key(t->c);
        seis3d::KernelConfig best{};
        double best_t = -1;
        {
            std::lock_guard<std::mutex> lock(g_tune_mutex);
            auto it = g_tuned.find(key);
            if (it != g_tuned.end()) { best = it->second.config; best_t = it->second.ms_per_step * steps; }
        }
        if (best_t < 0) {
            // Time every candidate on a scratch session of the same grid, so the
            // caller's wavefield and model are untouched.
            seis3d_t* scratch = nullptr;
            const int rc = seis3d_create(&t->c, &scratch);
            if (rc != SEIS3D_OK) fail(rc, g_error);
            try {
                hipEvent_t e0 = nullptr, e1 = nullptr;
                check(hipEventCreate(&e0), "create event");
                check(hipEventCreate(&e1), "create event");
                for (const auto& cand : tile_candidates(t->c)) {
                    scratch->kernels = cand;
                    enqueue_step(scratch, *src, 0);   // warm-up
                    check(hipEventRecord(e0, scratch->stream), "record event");
                    for (int k = 0; k < steps; ++k) enqueue_step(scratch, *src, std::uint64_t(k + 1));
                    check(hipEventRecord(e1, scratch->stream), "record event");
                    check(hipEventSynchronize(e1), "autotune");
                    float ms = 0;
                    check(hipEventElapsedTime(&ms, e0, e1), "read event");
                    if (best_t < 0 || ms < best_t) { best_t = ms; best = cand; }
                }
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
            } catch (...) {
                seis3d_destroy(scratch);
                throw;
            }
            seis3d_destroy(scratch);
            std::lock_guard<std::mutex> lock(g_tune_mutex);
            g_tuned[key] = {best, best_t / steps};
        }
        t->kernels = best;
        if (family) *family = best.family;
        if (tile_y) *tile_y = best.tile_y;
        if (z_chunk) *z_chunk = best.z_chunk;
        if (best_ms) *best_ms = best_t / steps;
    });
}

}  // extern "C"

namespace {

// Shots with CPML or a free top (no reflecting walls, no learned frame) on one device step in
// multi-shot launches: each half-step of every shot is one launch per tile height.
bool multi_shot_ok(const std::vector<seis3d_t*>& sessions) {
    for (auto* t : sessions)
        if (t->c.reflecting || t->abc || t->device != sessions[0]->device) return false;
    return true;
}

// The LDS-DMA shape a shot steps with in a multi-shot launch: its own when it uses that family
// with uniform chunks; the default shape for a shot on the direct family or on the balanced
// one-round launch (z_chunk 0), which is a single-shot launch (every shape rounds identically).
seis3d::KernelConfig multi_shot_shape(const seis3d_t* t) {
    seis3d::KernelConfig k = t->kernels;
    if (k.family != 4 || (k.tile_y != 4 && k.tile_y != 8) || k.z_chunk < 1 || k.z_chunk > 128) {
        k.family = 4;
        k.tile_y = 4;
        k.z_chunk = std::min(16, std::max(1, t->c.gz - 8));
    }
    return k;
}

// Builds the multi-shot tables and captures every step of every shot.  Shots are grouped by
// tile height, at most kShotsPerLaunch per group; each group steps with one launch per
// half-step on its own stream (the stream of its first session), forked from and joined to
// b->origin, so that different groups' velocity and stress launches overlap.  The gather
// bookkeeping advances as maybe_sample would.
void capture_multi_shot(seis3d_batch* b, const seis3d_source* sources, int64_t n_steps) {
    const int S = int(b->sessions.size());
    std::vector<int> order(S);
    for (int s = 0; s < S; ++s) order[s] = s;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
        return multi_shot_shape(b->sessions[a]).tile_y < multi_shot_shape(b->sessions[c]).tile_y;
    });
    struct Group {
        int begin, end, tile_y;
        int max_gx, max_gy, max_rec;
        bool any_free;
        std::vector<std::size_t> start;   // per step: first sample entry of this group
    };
    std::vector<Group> groups;
    std::vector<seis3d::DevStep> table(std::size_t(2 * n_steps) * S);
    std::vector<seis3d::TileMap> maps(S);
    std::vector<float> spacing(S);
    for (int j = 0; j < S; ++j) {
        seis3d_t* t = b->sessions[order[j]];
        const seis3d::KernelConfig k = multi_shot_shape(t);
        if (groups.empty() || groups.back().tile_y != k.tile_y ||
            groups.back().end - groups.back().begin == seis3d::kShotsPerLaunch)
            groups.push_back({j, j, k.tile_y, 8, 8, 0, false, {}});
        Group& gr = groups.back();
        gr.end = j + 1;
        gr.max_gx = std::max(gr.max_gx, t->c.gx);
        gr.max_gy = std::max(gr.max_gy, t->c.gy);
        gr.max_rec = std::max(gr.max_rec, t->receiver_count);
        gr.any_free = gr.any_free || t->c.free_surface;
        maps[j] = seis3d::dma_tile_map(t->base().g, k.tile_y, k.z_chunk);
        spacing[j] = t->c.spacing;
    }
    // The kernel argument of one group's launch for table row `row` (steps of one half-step).
    auto launch_args = [&](const Group& gr, const seis3d::DevStep* row) {
        seis3d::ShotLaunch L{};
        L.nshots = gr.end - gr.begin;
        for (int i = 0; i < L.nshots; ++i) {
            L.maps[i] = maps[gr.begin + i];
            L.steps[i] = row[gr.begin + i];
            L.first[i + 1] = L.first[i] + L.maps[i].nblocks;
        }
        L.nblocks = L.first[L.nshots];
        L.per_xcd = (L.nblocks + 7) / 8;
        return L;
    };
    for (int64_t n = 0; n < n_steps; ++n) {
        for (int j = 0; j < S; ++j) {
            const int s = order[j];
            seis3d_t* t = b->sessions[s];
            const seis3d_ctx& c = t->c;
            seis3d::DevStep V = t->base(), St = V;
            if (sources[s].kind == SEIS3D_FORCE) V.src = make_source(c, sources[s], double(n) * c.dt);
            if (sources[s].kind == SEIS3D_EXPLOSION) St.src = make_source(c, sources[s], (double(n) + 0.5) * c.dt);
            table[std::size_t(2 * n) * S + j] = V;
            table[std::size_t(2 * n + 1) * S + j] = St;
        }
    }
    // Receiver samples, per group and step; entries index the group's rows of the step table.
    std::vector<seis3d::ShotSample> samples;
    for (auto& gr : groups) {
        gr.start.assign(std::size_t(n_steps) + 1, samples.size());
        for (int64_t n = 0; n < n_steps; ++n) {
            for (int j = gr.begin; j < gr.end; ++j) {
                seis3d_t* t = b->sessions[order[j]];
                if (t->recording && t->gather_count < t->gather_target &&
                    (std::uint64_t(n) + 1) % t->gather_stride == 0) {
                    samples.push_back({j - gr.begin, t->receiver_count, t->receivers,
                                       t->gather + std::size_t(t->gather_count) * t->gather_receivers * 3});
                    if (++t->gather_count >= t->gather_target) t->recording = false;
                }
            }
            gr.start[std::size_t(n) + 1] = samples.size();
        }
    }
    check(hipMalloc(&b->d_steps, table.size() * sizeof(seis3d::DevStep)), "allocate batch tables");
    check(hipMalloc(&b->d_spacing, spacing.size() * sizeof(float)), "allocate batch tables");
    if (!samples.empty())
        check(hipMalloc(&b->d_samples, samples.size() * sizeof(seis3d::ShotSample)), "allocate batch tables");
    // The tables go up before capture starts (a synchronous copy may not be captured).
    check(hipMemcpy(b->d_steps, table.data(), table.size() * sizeof(seis3d::DevStep), hipMemcpyHostToDevice),
          "upload batch tables");
    check(hipMemcpy(b->d_spacing, spacing.data(), spacing.size() * sizeof(float), hipMemcpyHostToDevice),
          "upload batch tables");
    if (!samples.empty())
        check(hipMemcpy(b->d_samples, samples.data(), samples.size() * sizeof(seis3d::ShotSample),
                        hipMemcpyHostToDevice),
              "upload batch tables");
    check(hipStreamBeginCapture(b->origin, hipStreamCaptureModeGlobal), "begin capture");
    check(hipEventRecord(b->events[0], b->origin), "fork capture");
    for (std::size_t q = 0; q < groups.size(); ++q) {
        const Group& gr = groups[q];
        hipStream_t st = b->sessions[order[gr.begin]]->stream;
        check(hipStreamWaitEvent(st, b->events[0], 0), "fork capture");
        for (int j = gr.begin; j < gr.end; ++j) {
            seis3d_t* t = b->sessions[order[j]];
            if (t->c.free_surface) check(seis3d::launch_stress_ghost(t->base(), st), "kernel launch");
        }
        for (int64_t n = 0; n < n_steps; ++n) {
            const seis3d::DevStep* hv = table.data() + std::size_t(2 * n) * S;
            const seis3d::DevStep* vel = b->d_steps + std::size_t(2 * n) * S + gr.begin;
            check(seis3d::launch_dma_shots(true, launch_args(gr, hv), gr.tile_y, st), "kernel launch");
            if (gr.any_free)
                check(seis3d::launch_velocity_ghost_shots(vel, b->d_spacing + gr.begin, gr.end - gr.begin, gr.max_gx,
                                                          gr.max_gy, st),
                      "kernel launch");
            check(seis3d::launch_dma_shots(false, launch_args(gr, hv + S), gr.tile_y, st), "kernel launch");
            const std::size_t k = gr.start[std::size_t(n) + 1] - gr.start[std::size_t(n)];
            if (k)
                check(seis3d::launch_receivers_shots(vel, b->d_samples + gr.start[std::size_t(n)], int(k), gr.max_rec, st),
                      "kernel launch");
        }
        check(hipEventRecord(b->events[q + 1], st), "join capture");
        check(hipStreamWaitEvent(b->origin, b->events[q + 1], 0), "join capture");
    }
    check(hipStreamEndCapture(b->origin, &b->graph), "end capture");
}

// Captures every step of every shot as one work-queue launch per half-step (plus one velocity-ghost
// and one receiver launch per step for all shots) on b->origin.  The step table holds each shot's
// parameters per half-step, as in the shared-launch mode; the gather bookkeeping advances as
// maybe_sample would.
void capture_queue(seis3d_batch* b, const seis3d_source* sources, int64_t n_steps) {
    const int S = int(b->sessions.size());
    std::vector<seis3d::DevStep> table(std::size_t(2 * n_steps) * S);
    std::vector<seis3d::TileMap> maps(S);
    std::vector<int> first(std::size_t(S) + 1, 0);
    std::vector<float> spacing(S);
    int max_gx = 8, max_gy = 8, max_rec = 0;
    bool any_free = false;
    for (int s = 0; s < S; ++s) {
        seis3d_t* t = b->sessions[s];
        // 64x4 tiles for every shot, with the shot's own (autotuned) z-chunk length.
        maps[s] = seis3d::dma_tile_map(t->base().g, 4, multi_shot_shape(t).z_chunk);
        first[std::size_t(s) + 1] = first[std::size_t(s)] + maps[s].nblocks;
        spacing[s] = t->c.spacing;
        max_gx = std::max(max_gx, t->c.gx);
        max_gy = std::max(max_gy, t->c.gy);
        max_rec = std::max(max_rec, t->receiver_count);
        any_free = any_free || t->c.free_surface;
    }
    const int total = first[std::size_t(S)];
    // The eight XCDs' ranges of the item list, of equal work: an item of shot s marches its
    // chunk's planes plus 8 halo planes (counted as half a plane each, the halo planes stage only
    // the tile's own rows).
    std::vector<int> xcd_first(9, 0);
    {
        std::vector<double> work(static_cast<std::size_t>(S));
        double all = 0;
        for (int s = 0; s < S; ++s) {
            const double per_item = double(b->sessions[s]->c.gz - 8) / maps[s].nbz + 4.0;
            work[std::size_t(s)] = per_item;
            all += per_item * maps[s].nblocks;
        }
        int x = 1;
        double acc = 0;
        for (int s = 0; s < S && x < 8; ++s) {
            for (int k = 0; k < maps[s].nblocks && x < 8; ++k) {
                acc += work[std::size_t(s)];
                while (x < 8 && acc >= all * x / 8) xcd_first[std::size_t(x++)] = first[std::size_t(s)] + k + 1;
            }
        }
        while (x < 8) xcd_first[std::size_t(x++)] = total;
        xcd_first[8] = total;
    }
    for (int64_t n = 0; n < n_steps; ++n) {
        for (int s = 0; s < S; ++s) {
            seis3d_t* t = b->sessions[s];
            const seis3d_ctx& c = t->c;
            seis3d::DevStep V = t->base(), St = V;
            if (sources[s].kind == SEIS3D_FORCE) V.src = make_source(c, sources[s], double(n) * c.dt);
            if (sources[s].kind == SEIS3D_EXPLOSION) St.src = make_source(c, sources[s], (double(n) + 0.5) * c.dt);
            table[std::size_t(2 * n) * S + s] = V;
            table[std::size_t(2 * n + 1) * S + s] = St;
        }
    }
    std::vector<seis3d::ShotSample> samples;
    std::vector<std::size_t> start(std::size_t(n_steps) + 1, 0);
    for (int64_t n = 0; n < n_steps; ++n) {
        for (int s = 0; s < S; ++s) {
            seis3d_t* t = b->sessions[s];
            if (t->recording && t->gather_count < t->gather_target && (std::uint64_t(n) + 1) % t->gather_stride == 0) {
                samples.push_back({s, t->receiver_count, t->receivers,
                                   t->gather + std::size_t(t->gather_count) * t->gather_receivers * 3});
                if (++t->gather_count >= t->gather_target) t->recording = false;
            }
        }
        start[std::size_t(n) + 1] = samples.size();
    }
    int blocks = 0;
    check(seis3d::queue_grid(&blocks), "size work-queue grid");
    const std::size_t ncounters = std::size_t(n_steps) * 16;
    check(hipMalloc(&b->d_steps, table.size() * sizeof(seis3d::DevStep)), "allocate batch tables");
    check(hipMalloc(&b->d_spacing, spacing.size() * sizeof(float)), "allocate batch tables");
    check(hipMalloc(&b->d_maps, maps.size() * sizeof(seis3d::TileMap)), "allocate batch tables");
    check(hipMalloc(&b->d_first, (first.size() + xcd_first.size()) * sizeof(int)), "allocate batch tables");
    check(hipMalloc(&b->d_counters, ncounters * sizeof(unsigned)), "allocate batch tables");
    if (!samples.empty())
        check(hipMalloc(&b->d_samples, samples.size() * sizeof(seis3d::ShotSample)), "allocate batch tables");
    // The tables go up before capture starts (a synchronous copy may not be captured).
    check(hipMemcpy(b->d_steps, table.data(), table.size() * sizeof(seis3d::DevStep), hipMemcpyHostToDevice),
          "upload batch tables");
    check(hipMemcpy(b->d_spacing, spacing.data(), spacing.size() * sizeof(float), hipMemcpyHostToDevice),
          "upload batch tables");
    check(hipMemcpy(b->d_maps, maps.data(), maps.size() * sizeof(seis3d::TileMap), hipMemcpyHostToDevice),
          "upload batch tables");
    check(hipMemcpy(b->d_first, first.data(), first.size() * sizeof(int), hipMemcpyHostToDevice),
          "upload batch tables");
    const int* d_xcd_first = b->d_first + first.size();
    check(hipMemcpy(b->d_first + first.size(), xcd_first.data(), xcd_first.size() * sizeof(int),
                    hipMemcpyHostToDevice),
          "upload batch tables");
    if (!samples.empty())
        check(hipMemcpy(b->d_samples, samples.data(), samples.size() * sizeof(seis3d::ShotSample),
                        hipMemcpyHostToDevice),
              "upload batch tables");
    hipStream_t st = b->origin;
    check(hipStreamBeginCapture(st, hipStreamCaptureModeGlobal), "begin capture");
    // Every replay starts with fresh queues.
    check(seis3d::launch_zero_words(b->d_counters, (long long)ncounters, st), "kernel launch");
    for (auto* t : b->sessions)
        if (t->c.free_surface) check(seis3d::launch_stress_ghost(t->base(), st), "kernel launch");
    for (int64_t n = 0; n < n_steps; ++n) {
        const seis3d::DevStep* vel = b->d_steps + std::size_t(2 * n) * S;
        unsigned* q = b->d_counters + std::size_t(n) * 16;
        check(seis3d::launch_dma_queue(true, vel, b->d_maps, b->d_first, d_xcd_first, q, S, blocks, st),
              "kernel launch");
        if (any_free)
            check(seis3d::launch_velocity_ghost_shots(vel, b->d_spacing, S, max_gx, max_gy, st), "kernel launch");
        check(seis3d::launch_dma_queue(false, vel + S, b->d_maps, b->d_first, d_xcd_first, q + 8, S, blocks, st),
              "kernel launch");
        const std::size_t k = start[std::size_t(n) + 1] - start[std::size_t(n)];
        if (k)
            check(seis3d::launch_receivers_shots(vel, b->d_samples + start[std::size_t(n)], int(k), max_rec, st),
                  "kernel launch");
    }
    check(hipStreamEndCapture(st, &b->graph), "end capture");
}

}  // namespace

extern "C" {

int seis3d_batch_create(seis3d_t* const* sessions, const seis3d_source* sources, int count, int64_t n_steps,
                        seis3d_batch_t** out) {
    return seis3d_batch_create_mode(sessions, sources, count, n_steps, SEIS3D_BATCH_SHOT_STREAMS, out);
}

int seis3d_batch_create_mode(seis3d_t* const* sessions, const seis3d_source* sources, int count, int64_t n_steps,
                             int mode, seis3d_batch_t** out) {
    return guarded([&] {
        if (mode != SEIS3D_BATCH_SHOT_STREAMS && mode != SEIS3D_BATCH_SHARED_LAUNCHES &&
            mode != SEIS3D_BATCH_WORK_QUEUE)
            invalid("Seis3D: unknown batch mode.");
        if (!sessions || !sources || !out || count < 1 || count > 256 || n_steps < 1 || n_steps > 1000000)
            invalid("Seis3D: invalid batch request.");
        for (int s = 0; s < count; ++s) {
            require_session(sessions[s]);
            check_source(sessions[s]->c, sources[s]);
            for (int r = 0; r < s; ++r)
                if (sessions[r] == sessions[s]) invalid("Seis3D: a session appears twice in one batch.");
        }
        auto* b = new seis3d_batch();
        b->sessions.assign(sessions, sessions + count);
        b->steps = n_steps;
        b->before.resize(std::size_t(count));
        b->after.resize(std::size_t(count));
        try {
            check(hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking), "create stream");
            b->events.resize(std::size_t(count) + 1, nullptr);
            for (auto& e : b->events) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "create event");
            // Work that depends on session state runs now, outside the graph: 1/mu after a
            // model change.  Model uploads are refused while the batch exists, so it stays valid.
            for (int s = 0; s < count; ++s) {
                seis3d_t* t = b->sessions[s];
                if (t->inv_mu_dirty) {
                    check(seis3d::launch_reciprocal_mu(t->base(), t->stream), "kernel launch");
                    t->inv_mu_dirty = false;
                }
                check(hipStreamSynchronize(t->stream), "batch capture");
                b->before[std::size_t(s)] = {t->gather_count, t->recording, t->ghosts_valid};
            }
            // Each shot starts by rebuilding its stress ghost planes, so a launch is valid
            // whatever was written into the wavefield since the last one.
            b->multi = mode != SEIS3D_BATCH_SHOT_STREAMS && multi_shot_ok(b->sessions);
            if (b->multi) {
                // One launch per half-step for a group of shots or for all of them: the shots'
                // tiles share the GPU.
                if (mode == SEIS3D_BATCH_WORK_QUEUE) capture_queue(b, sources, n_steps);
                else capture_multi_shot(b, sources, n_steps);
                for (int s = 0; s < count; ++s) {
                    seis3d_t* t = b->sessions[s];
                    b->after[std::size_t(s)] = {t->gather_count, t->recording, true};
                }
            } else {
                // Every session's stream joins one capture: fork from the origin stream,
                // enqueue all steps of every shot on its own stream, join back.
                check(hipStreamBeginCapture(b->origin, hipStreamCaptureModeGlobal), "begin capture");
                check(hipEventRecord(b->events[0], b->origin), "fork capture");
                for (int s = 0; s < count; ++s) {
                    seis3d_t* t = b->sessions[s];
                    check(hipStreamWaitEvent(t->stream, b->events[0], 0), "fork capture");
                    if (t->c.free_surface) check(seis3d::launch_stress_ghost(t->base(), t->stream), "kernel launch");
                    t->ghosts_valid = true;
                    for (int64_t k = 0; k < n_steps; ++k) {
                        enqueue_step(t, sources[s], std::uint64_t(k));
                        maybe_sample(t, std::uint64_t(k));
                    }
                    b->after[std::size_t(s)] = {t->gather_count, t->recording, true};
                    check(hipEventRecord(b->events[s + 1], t->stream), "join capture");
                    check(hipStreamWaitEvent(b->origin, b->events[s + 1], 0), "join capture");
                }
                check(hipStreamEndCapture(b->origin, &b->graph), "end capture");
            }
            check(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "instantiate graph");
        } catch (...) {
            if (b->origin) {
                hipGraph_t partial = nullptr;
                (void)hipStreamEndCapture(b->origin, &partial);
                if (partial) (void)hipGraphDestroy(partial);
            }
            // Capture only recorded work: the gather bookkeeping is as it was.
            for (std::size_t s = 0; s < b->sessions.size(); ++s) {
                b->sessions[s]->gather_count = b->before[s].count;
                b->sessions[s]->recording = b->before[s].recording;
                b->sessions[s]->ghosts_valid = b->before[s].ghosts;
            }
            b->release();
            delete b;
            throw;
        }
        // Nothing has run yet: samples are counted when the batch is launched.
        for (std::size_t s = 0; s < b->sessions.size(); ++s) {
            b->sessions[s]->gather_count = b->before[s].count;
            b->sessions[s]->recording = b->before[s].recording;
            b->sessions[s]->ghosts_valid = b->before[s].ghosts;
            ++b->sessions[s]->batch_refs;
        }
        b->holds = true;
        *out = b;
    });
}

int seis3d_batch_launch(seis3d_batch_t* b) {
    return guarded([&] {
        if (!b || !b->exec) fail(SEIS3D_ELIFECYCLE, "Seis3D: batch not created.");
        for (auto* t : b->sessions)
            if (t->destroy_pending) fail(SEIS3D_ELIFECYCLE, "Seis3D: a session of this batch was destroyed.");
        check(hipGraphLaunch(b->exec, b->origin), "launch batch");
        for (std::size_t s = 0; s < b->sessions.size(); ++s) {
            b->sessions[s]->gather_count = b->after[s].count;
            b->sessions[s]->recording = b->after[s].recording;
            b->sessions[s]->ghosts_valid = true;
        }
    });
}

int seis3d_batch_synchronize(seis3d_batch_t* b) {
    return guarded([&] {
        if (!b || !b->origin) fail(SEIS3D_ELIFECYCLE, "Seis3D: batch not created.");
        check(hipStreamSynchronize(b->origin), "finish batch");
    });
}

void seis3d_batch_destroy(seis3d_batch_t* b) {
    if (!b) return;
    b->release();
    delete b;
}

}  // extern "C"
