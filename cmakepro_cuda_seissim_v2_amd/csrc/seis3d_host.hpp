// seis3d_host.hpp - internal host-side declarations shared by the C ABI
// (seis3d_api.cpp) and the adjoint / FWI driver (seis3d_fwi.cpp): error
// plumbing, the device allocation plan and the session object.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/seis3d.h"
#include "seis3d_kernels.hpp"

namespace seis3d::host {


inline constexpr int kRadius = 4;
inline constexpr std::size_t kMaxPhysicalNodes = std::size_t(300) * 300 * 400;   // Common3D.h:10
inline constexpr int kMaxVolumeAxis = 256;                                        // Common3D.h:11
inline constexpr int kMaxVolumeOutput = 1536;                                     // Common3D.h:12
inline constexpr int kMaxReceivers = 300;                                         // Common3D.h:14
inline constexpr int kMaxSegySamples = 32767;                                     // Common3D.h:15
inline constexpr int kMaxSegyIntervalUs = 32767;                                  // Common3D.h:16

// Grid arrays start kGridLead floats into their allocation, so that the first
// interior column (x = 4) of every row is 128-byte aligned: row pitches are
// multiples of 64 floats, and 28 + 4 = 32 floats = 128 bytes.  A 64-column
// tile row is then exactly two cache lines.
inline constexpr std::size_t kGridLead = 28;
inline void free_grid(float*& p) noexcept {
    if (p) (void)hipFree(p - kGridLead);
    p = nullptr;
}
inline constexpr float kSourceNorm = 10.8387785335f;                              // (1 + 2 exp(-1/2))^3

inline thread_local std::string g_error;

struct Failure {
    int code;
    std::string message;
};

[[noreturn]] inline void fail(int code, const std::string& message) { throw Failure{code, message}; }
[[noreturn]] inline void invalid(const std::string& message) { fail(SEIS3D_EINVAL, message); }

inline void check(hipError_t e, const char* operation) {
    if (e != hipSuccess) fail(SEIS3D_EHIP, std::string("Seis3D / ") + operation + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F&& fn) {
    try {
        fn();
        return SEIS3D_OK;
    } catch (const Failure& f) {
        g_error = f.message;
        return f.code;
    } catch (const std::bad_alloc&) {
        g_error = "Seis3D: host allocation failed.";
        return SEIS3D_EHIP;
    } catch (const std::exception& e) {
        g_error = e.what();
        return SEIS3D_EHIP;
    }
}

inline int active_pml(const seis3d_ctx& c) { return c.reflecting ? 0 : c.pml; }
inline std::size_t align_up(std::size_t v, std::size_t a) { return (v + a - 1) / a * a; }

// Device allocation plan for one context.
struct Plan {
    int px = 0;
    std::size_t plane = 0, cells = 0;   // pitched cells
    std::size_t slab[3]{};              // floats per CPML slot
    std::size_t bytes = 0;
};

// validateContext3D: checks the context and derives the padded shape and dt.
void validate(seis3d_ctx& c);
// One material cell: finite, 0 < Vp <= 30000, 0 <= Vs < Vp, positive bulk modulus, 0 < rho <= 50000.
bool valid_material(float vp, float vs, float rho);
Plan plan_for(const seis3d_ctx& c);
// CPML tables [half][n][a, b, 1/kappa] of one axis: the reference profile, or with `profile`
// (SEIS3D_PROFILE_PARAMS values, checked by check_profile) the learned one.
std::vector<float> make_coefficients(const seis3d_ctx& c, int axis, const double* profile = nullptr);
void check_profile(const double* profile);
float ricker(float frequency, double time);

}  // namespace seis3d::host

// ---------------------------------------------------------------------------
// Session
// ---------------------------------------------------------------------------

namespace seis3d::host {
struct AbcFrame;                                   // learned-boundary state (seis3d_abc.cpp)
void release_abc(AbcFrame*& a) noexcept;
}  // namespace seis3d::host

struct seis3d_session {
    using Plan = seis3d::host::Plan;
    seis3d_ctx c{};
    Plan plan;
    int device = 0;
    hipStream_t stream = nullptr;
    float* fields[9]{};
    float* inv_rho = nullptr;
    float* lambda = nullptr;
    float* mu = nullptr;
    float* inv_mu = nullptr;           // 1/mu, rebuilt before the next step after a model change
    bool inv_mu_dirty = true;
    float* psi[3]{};
    float* coef[3]{};
    unsigned int* flags = nullptr;
    int* receivers = nullptr;          // padded coordinates, 3 ints each
    float* receiver_values = nullptr;
    float* host_values = nullptr;      // pinned
    int receiver_count = 0;
    float* gather = nullptr;           // [sample][receiver][3]
    float* stage = nullptr;            // gx*gy*gz floats: field transfers, allocated on first use
    int gather_target = 0, gather_count = 0, gather_stride = 1, gather_receivers = 0;
    bool recording = false;
    bool ghosts_valid = true;          // σ ghost planes consistent with the interior
    seis3d::KernelConfig kernels;
    seis3d::host::AbcFrame* abc = nullptr;   // learned absorbing boundary, when attached
    bool has_profile = false;                // CPML tables follow `profile`, not the reference's
    double profile[SEIS3D_PROFILE_PARAMS]{};
    const double* cpml_profile() const { return has_profile ? profile : nullptr; }
    // Batches whose captured graph holds this session's device pointers.  While any
    // exists, calls that free or replace those buffers are refused, and a destroy is
    // deferred until the last such batch is destroyed.
    int batch_refs = 0;
    bool destroy_pending = false;

    seis3d::DevStep base() const {
        seis3d::DevStep P{};
        P.g.gx = c.gx; P.g.gy = c.gy; P.g.gz = c.gz;
        P.g.px = plan.px;
        P.g.plane = static_cast<long long>(plan.plane);
        P.g.pml = seis3d::host::active_pml(c);
        P.g.free_surface = c.free_surface;
        P.g.tz = (c.free_surface ? 1 : 2) * P.g.pml;
        for (int a = 0; a < 3; ++a) P.g.slab[a] = static_cast<long long>(plan.slab[a]);
        for (int f = 0; f < 9; ++f) P.p.f[f] = P.p.out[f] = fields[f];
        P.p.inv_rho = inv_rho; P.p.lambda = lambda; P.p.mu = mu; P.p.inv_mu = inv_mu;
        for (int a = 0; a < 3; ++a) { P.p.psi[a] = P.p.psi_out[a] = psi[a]; P.p.coef[a] = coef[a]; }
        P.p.flags = flags;
        P.dt = c.dt;
        P.invh = 1.0f / c.spacing;
        P.src.active = 0;
        return P;
    }

    seis3d::MaterialParams material_params() const {
        seis3d::MaterialParams m{};
        m.nx = c.nx; m.ny = c.ny; m.nz = c.nz; m.ox = c.ox; m.oy = c.oy; m.oz = c.oz;
        m.kind = c.model_kind;
        m.vp = c.vp; m.vs = c.vs; m.rho = c.density;
        m.layer_vp = c.layer_vp; m.layer_vs = c.layer_vs; m.layer_rho = c.layer_density;
        m.interface_depth = c.interface_depth; m.layer_dip = c.layer_dip; m.layer_thickness = c.layer_thickness;
        m.fault_throw = c.fault_throw; m.anomaly_radius = c.anomaly_radius;
        return m;
    }

    void release() noexcept {
        if (stream) (void)hipStreamSynchronize(stream);
        seis3d::host::release_abc(abc);
        for (auto& p : fields) seis3d::host::free_grid(p);
        seis3d::host::free_grid(inv_rho); seis3d::host::free_grid(lambda);
        seis3d::host::free_grid(mu); seis3d::host::free_grid(inv_mu);
        for (int a = 0; a < 3; ++a) { (void)hipFree(psi[a]); (void)hipFree(coef[a]); psi[a] = coef[a] = nullptr; }
        (void)hipFree(flags); (void)hipFree(receivers); (void)hipFree(receiver_values); (void)hipFree(gather);
        (void)hipFree(stage);
        flags = nullptr; receivers = nullptr; receiver_values = nullptr; gather = nullptr; stage = nullptr;
        if (host_values) (void)hipHostFree(host_values);
        host_values = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

namespace seis3d::host {

void require_session(const seis3d_t* t);
// Refuses (SEIS3D_ELIFECYCLE) an operation that would free or replace device buffers a batch holds.
void require_unbatched(const seis3d_t* t, const char* operation);
void alloc_floats(float*& p, std::size_t count, hipStream_t s, const char* what);
// A zeroed grid array of `cells` floats at kGridLead into its allocation (free with free_grid).
void alloc_grid(float*& p, std::size_t cells, hipStream_t s, const char* what);
void check_source(const seis3d_ctx& c, const seis3d_source& s);
// Source term at time `time` (sourceKernel3D, cuda/Cuda_Seis3D.cu:313-332).
seis3d::DevSource make_source(const seis3d_ctx& c, const seis3d_source& s, double time);
// One runStep3D (cuda/Cuda_Seis3D.cu:602-624), enqueued on the session stream;
// optional events bracket the two fused half-steps.
void enqueue_step(seis3d_t* t, const seis3d_source& s, std::uint64_t n, hipEvent_t* ev = nullptr);
// Learned-boundary correction of the frame cells after the stress half-step (no-op when detached).
void enqueue_abc(seis3d_t* t);
// Samples of the packed views (planes, plus the volume when `volume`), 2 floats each.
long long view_samples(const seis3d_ctx& c, int volume);
// extractViews3D through caller-provided device staging (2 floats per sample) and 5 statistics
// words; validates the request and synchronises the session stream.
void extract_views_staged(seis3d_t* t, float* out, size_t capacity, const int32_t slices[3], int component,
                          int material_component, int volume, int include_boundary, seis3d_view_stats* stats,
                          float* stage, unsigned* stat_words);

}  // namespace seis3d::host
