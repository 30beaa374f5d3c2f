// seis3d_oracle.hpp - CPU restatement of the reference 3D O(8) velocity-stress
// scheme, templated on the arithmetic type (float = the reference FP32 path,
// double = an FP64 run of the same scheme).
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
// check of __graft_entry__.py and the cpu_baseline leg of bench.py may build,
// load or call it.  The product (cmakepro_cuda_seissim_v2_amd/) never does.
//
// Every function cites the reference file:line it restates.  The FP32
// instantiation keeps the reference's expression trees operation for
// operation; it is meant to be compiled with -ffp-contract=off so that no
// multiply-add is fused (the HIP kernels are built the same way).
//
// Pinned by (tests/test_oracle3d.py): O(8) moment conditions, the CFL
// ceiling and integer-microsecond dt closed form, CPML coefficients against
// the closed-form formula and a fresh-vs-refresh bitwise check, a cubic
// stress field giving the analytic velocity, free-surface ghost parity,
// source weights summing to one, receiver layout formulas, plane-wave
// travel time, FP32-vs-FP64 agreement, CPML late/peak decay and the
// reflecting-wall contrast of tests/Seis3DBoundaryRegression.cpp:89.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../include/seis3d.h"

namespace seis3d_oracle {

constexpr int kRadius = 4;                                      // Common3D.h:9
constexpr std::size_t kMaxPhysicalNodes = std::size_t(300) * 300 * 400;  // Common3D.h:10
constexpr int kMaxReceivers = 300;                              // Common3D.h:14
constexpr int kMaxVolumeAxis = 256;                             // Common3D.h:11
constexpr int kMaxVolumeOutput = 1536;                          // Common3D.h:12

struct Invalid : std::invalid_argument { using std::invalid_argument::invalid_argument; };

// validMaterial3D (cuda/Cuda_Seis3D.cu:30-34).
inline bool valid_material(float vp, float vs, float rho) {
    return std::isfinite(vp) && std::isfinite(vs) && std::isfinite(rho)
        && vp > 0 && vp <= 30000 && vs >= 0 && vs < vp
        && vp * vp > (4.0f / 3.0f) * vs * vs && rho > 0 && rho <= 50000;
}

inline int active_pml(const seis3d_ctx& c) { return c.reflecting ? 0 : c.pml; }  // Common3D.h:63-65

// validateContext3D (cuda/Cuda_Seis3D.cu:433-483).
inline void validate(seis3d_ctx& c) {
    if (c.reflecting) c.free_surface = 0;
    if (c.model_kind < 0 || c.model_kind > SEIS3D_MODEL_SALT_DOME
        || !std::isfinite(c.interface_depth) || c.interface_depth < .1f || c.interface_depth > .9f
        || !std::isfinite(c.layer_dip) || std::abs(c.layer_dip) > .6f
        || !std::isfinite(c.layer_thickness) || c.layer_thickness < .03f || c.layer_thickness > .5f
        || !std::isfinite(c.fault_throw) || std::abs(c.fault_throw) > .4f
        || !std::isfinite(c.anomaly_radius) || c.anomaly_radius < .05f || c.anomaly_radius > .4f)
        throw Invalid("Seis3D: invalid procedural model geometry.");
    if (c.nx < 16 || c.ny < 16 || c.nz < 16)
        throw Invalid("Seis3D: every physical dimension must be at least 16.");
    if (std::size_t(c.nx) > kMaxPhysicalNodes / std::size_t(c.ny)
        || std::size_t(c.nx) * c.ny > kMaxPhysicalNodes / std::size_t(c.nz))
        throw Invalid("Seis3D: physical node count exceeds 300 x 300 x 400 (36,000,000).");
    if (c.pml < 8 || c.pml > 40 || !std::isfinite(c.spacing) || c.spacing <= 0
        || !std::isfinite(c.frequency) || c.frequency <= 0 || c.frequency > 10000
        || !std::isfinite(c.cfl) || c.cfl <= 0 || c.cfl > 0.9f)
        throw Invalid("Seis3D: invalid CPML, spacing, frequency or CFL safety factor.");
    if (!valid_material(c.vp, c.vs, c.density) || !valid_material(c.layer_vp, c.layer_vs, c.layer_density))
        throw Invalid("Seis3D: Vp/Vs/rho must be finite with positive bulk modulus.");
    if (c.model_kind != SEIS3D_MODEL_IMPORTED)
        c.max_vp = c.model_kind == SEIS3D_MODEL_UNIFORM ? c.vp : std::max(c.vp, c.layer_vp);
    if (!std::isfinite(c.max_vp) || c.max_vp <= 0 || c.max_vp > 30000)
        throw Invalid("Seis3D: invalid model maximum Vp.");
    const int pml = active_pml(c);
    c.ox = c.oy = pml + 4; c.oz = c.free_surface ? 4 : pml + 4;
    c.gx = c.nx + 2 * pml + 8; c.gy = c.ny + 2 * pml + 8;
    c.gz = c.nz + (c.free_surface ? 1 : 2) * pml + 8;
    c.cells = std::uint64_t(c.gx) * c.gy * c.gz;
    constexpr double sumWeights = 1225.0 / 1024 + 245.0 / 3072 + 49.0 / 5120 + 5.0 / 7168;
    const double safeDt = c.cfl * c.spacing / (std::sqrt(3.0) * c.max_vp * sumWeights);
    const auto microseconds = std::floor(safeDt * 1.0e6);
    if (microseconds < 1 || microseconds > 65535)
        throw Invalid("Seis3D: safe dt is outside the implemented 1..65535 microsecond range.");
    c.safe_dt = float(microseconds * 1.0e-6);
    c.dt = c.safe_dt;
    if (c.manual_dt) {
        const double requestedUs = double(c.requested_dt_ms) * 1000;
        const double roundedUs = std::round(requestedUs);
        if (!std::isfinite(requestedUs) || roundedUs < 1 || roundedUs > 65535)
            throw Invalid("Seis3D: manual dt must be between 0.001 and 65.535 ms.");
        if (std::abs(requestedUs - roundedUs) > 0.05)
            throw Invalid("Seis3D: manual dt must use whole microseconds (0.001 ms increments).");
        if (roundedUs > microseconds)
            throw Invalid("Seis3D: manual dt exceeds the model CFL ceiling; reduce dt or select Automatic CFL.");
        c.dt = float(roundedUs * 1.0e-6);
    }
}

// slabCount3D (cuda/Cuda_Seis3D.cu:37-42).
inline std::size_t slab_count(const seis3d_ctx& c, int axis) {
    if (c.reflecting) return 0;
    if (axis == 0) return std::size_t(2 * c.pml) * c.gy * c.gz;
    if (axis == 1) return std::size_t(2 * c.pml) * c.gx * c.gz;
    return std::size_t((c.free_surface ? 1 : 2) * c.pml) * c.gx * c.gy;
}

// makeCoefficients3D (cuda/Cuda_Seis3D.cu:45-64): FP64 profile, stored as Real
// ([half][n][a, b, 1/kappa]).  Real=float reproduces the reference's rounding.
template <class Real>
std::vector<Real> make_coefficients(const seis3d_ctx& c, int axis) {
    const int n = axis == 0 ? c.gx : axis == 1 ? c.gy : c.gz;
    std::vector<Real> values(std::size_t(n) * 6);
    const double sigmaMax = -4.0 * c.max_vp * std::log(1.0e-6) / (2.0 * c.pml * c.spacing);
    for (int half = 0; half < 2; ++half) for (int i = 0; i < n; ++i) {
        const double coordinate = i + 0.5 * half;
        double depth = (coordinate - (n - kRadius - c.pml - 0.5)) / c.pml;
        if (!(axis == 2 && c.free_surface))
            depth = std::max(depth, (kRadius + c.pml - 0.5 - coordinate) / c.pml);
        depth = std::clamp(depth, 0.0, 1.0);
        const double power = depth * depth * depth;
        const double sigma = sigmaMax * power, kappa = 1.0 + 2.0 * power;
        const double alpha = depth > 0 ? 3.141592653589793 * c.frequency * (1.0 - depth) : 0.0;
        const double b = std::exp(-(sigma / kappa + alpha) * c.dt);
        const double a = sigma > 0 ? sigma * (b - 1.0) / (kappa * (sigma + kappa * alpha)) : 0.0;
        const std::size_t offset = (std::size_t(half) * n + i) * 3;
        values[offset] = Real(a); values[offset + 1] = Real(b); values[offset + 2] = Real(1.0 / kappa);
    }
    return values;
}

// A learned CPML profile p = {log sigma scale, power, kappa max, alpha scale, alpha power} in the
// reference's table (no counterpart in the reference; its own profile is p = {0, 3, 3, 1, 1}):
// the layer depth of makeCoefficients3D (cuda/Cuda_Seis3D.cu:49-53) and its a/b formulas
// (:57-58), with sigma = e^p0 sigma_max depth^p1, kappa = 1 + (p2 - 1) depth^p1 and
// alpha = p3 pi f (1 - depth)^p4 for depth > 0.
template <class Real>
std::vector<Real> make_profile_coefficients(const seis3d_ctx& c, int axis, const double* p) {
    const int n = axis == 0 ? c.gx : axis == 1 ? c.gy : c.gz;
    std::vector<Real> table(std::size_t(n) * 6);
    const double pi = 3.141592653589793;
    const double reference_sigma = -4.0 * c.max_vp * std::log(1.0e-6) / (2.0 * c.pml * c.spacing);
    const double top = std::exp(p[0]) * reference_sigma;
    for (int half = 0; half < 2; ++half)
        for (int i = 0; i < n; ++i) {
            // Distance into the layer in layer widths, from whichever side is closer.
            const double x = i + 0.5 * half;
            double d = (x - (n - kRadius - c.pml - 0.5)) / c.pml;
            if (!(axis == 2 && c.free_surface)) d = std::max(d, (kRadius + c.pml - 0.5 - x) / c.pml);
            d = std::clamp(d, 0.0, 1.0);
            const double shape = std::pow(d, p[1]);
            const double sigma = top * shape;
            const double kappa = 1.0 + (p[2] - 1.0) * shape;
            const double alpha = d > 0 ? p[3] * pi * c.frequency * std::pow(1.0 - d, p[4]) : 0.0;
            const double b = std::exp(-(sigma / kappa + alpha) * c.dt);
            const double a = sigma > 0 ? sigma * (b - 1.0) / (kappa * (sigma + kappa * alpha)) : 0.0;
            Real* row = table.data() + (std::size_t(half) * n + i) * 3;
            row[0] = Real(a); row[1] = Real(b); row[2] = Real(1.0 / kappa);
        }
    return table;
}

// sourceKernel3D time function (cuda/Cuda_Seis3D.cu:319-320), FP64 phase.
inline double ricker_double(double frequency, double time) {
    const double phase = 3.141592653589793 * frequency * (time - 1.2 / frequency);
    return (1 - 2 * phase * phase) * std::exp(-phase * phase);
}

template <class Real> struct Weights;
// weight3D (cuda/Cuda_Seis3D.cu:72-75).
template <> struct Weights<float> {
    static float w(int r) {
        return r == 0 ? 1225.0f / 1024.0f : r == 1 ? -245.0f / 3072.0f : r == 2 ? 49.0f / 5120.0f : -5.0f / 7168.0f;
    }
};
template <> struct Weights<double> {
    static double w(int r) {
        return r == 0 ? 1225.0 / 1024.0 : r == 1 ? -245.0 / 3072.0 : r == 2 ? 49.0 / 5120.0 : -5.0 / 7168.0;
    }
};

template <class Real>
class Solver {
public:
    seis3d_ctx c{};
    std::vector<Real> f[9], inv_rho, lambda, mu;
    std::vector<Real> cpml[3];      // reference slab layout, [slot*slabCells + slab]
    std::size_t slab_cells[3]{};
    std::vector<Real> coef[3];      // [half][n][a,b,1/kappa]
    std::vector<seis3d_receiver> receivers;
    bool nonfinite = false;
    bool learned = false;           // CPML tables from `profile` (set_profile)
    double profile[5]{};

    // initializeGpu3D (cuda/Cuda_Seis3D.cu:519-555) minus device work: all
    // fields zero, CPML tables uploaded, procedural materials generated.
    explicit Solver(const seis3d_ctx& ctx) : c(ctx) {
        for (auto& v : f) v.assign(c.cells, Real(0));
        inv_rho.assign(c.cells, Real(0)); lambda.assign(c.cells, Real(0)); mu.assign(c.cells, Real(0));
        for (int a = 0; a < 3 && !c.reflecting; ++a) {
            slab_cells[a] = slab_count(c, a);
            cpml[a].assign(slab_cells[a] * 6, Real(0));
            coef[a] = make_coefficients<Real>(c, a);
        }
        generate_materials();
    }

    std::size_t index(int x, int y, int z) const {          // Common3D.h:96-98
        return (std::size_t(z) * c.gy + y) * c.gx + x;
    }

    // resetGpu3D (cuda/Cuda_Seis3D.cu:572-577).
    void reset() {
        for (auto& v : f) std::fill(v.begin(), v.end(), Real(0));
        for (auto& v : cpml) std::fill(v.begin(), v.end(), Real(0));
        nonfinite = false;
    }

    // refreshCpmlCoefficients3D (cuda/Cuda_Seis3D.cu:580-599); a learned profile is kept.
    void refresh_cpml(const seis3d_ctx& ctx) {
        if (c.reflecting) return;
        for (int a = 0; a < 3; ++a)
            coef[a] = learned ? make_profile_coefficients<Real>(ctx, a, profile) : make_coefficients<Real>(ctx, a);
        c = ctx;
    }

    // Tables of a learned CPML profile (nullptr: the reference's).
    void set_profile(const double* p) {
        learned = p != nullptr;
        if (p) std::copy(p, p + 5, profile);
        if (!c.reflecting) refresh_cpml(c);
    }

    // generateMaterialsKernel3D (cuda/Cuda_Seis3D.cu:114-147).  The blend is
    // evaluated in float as in the reference; the moduli in Real.
    void generate_materials() {
        for (std::size_t i = 0; i < c.cells; ++i) {
            const int x = std::max(0, std::min(c.nx - 1, int(i % c.gx) - c.ox));
            const int y = std::max(0, std::min(c.ny - 1, int((i / c.gx) % c.gy) - c.oy));
            const int z = std::max(0, std::min(c.nz - 1, int(i / (std::size_t(c.gx) * c.gy)) - c.oz));
            const float dx = (x - c.nx * 0.5f) / c.nx;
            const float dy = (y - c.ny * 0.5f) / c.ny;
            const float dz = (z - c.nz * 0.55f) / c.nz;
            const float depth = float(z) / float(c.nz - 1), horizontal = float(x) / float(c.nx - 1) - 0.5f;
            float blend = 0;
            switch (c.model_kind) {
            case SEIS3D_MODEL_TWO_LAYER: blend = depth >= c.interface_depth; break;
            case SEIS3D_MODEL_SPHERE: blend = dx * dx + dy * dy + dz * dz < c.anomaly_radius * c.anomaly_radius; break;
            case SEIS3D_MODEL_THREE_LAYER: blend = depth < 1.0f / 3 ? 0 : depth < 2.0f / 3 ? 0.5f : 1; break;
            case SEIS3D_MODEL_DIPPING_LAYER: blend = depth >= c.interface_depth + c.layer_dip * horizontal; break;
            case SEIS3D_MODEL_LOW_VELOCITY_LAYER: blend = std::fabs(depth - c.interface_depth) <= 0.5f * c.layer_thickness; break;
            case SEIS3D_MODEL_FAULT: blend = depth >= c.interface_depth + c.layer_dip * horizontal
                + (horizontal > 0 ? c.fault_throw : 0); break;
            case SEIS3D_MODEL_GRADIENT: blend = depth; break;
            case SEIS3D_MODEL_SALT_DOME: {
                const float vertical = (depth - 0.78f) / 0.52f;
                blend = (dx * dx + dy * dy) / (c.anomaly_radius * c.anomaly_radius) + vertical * vertical < 1;
                break;
            }
            default: break;
            }
            const Real vp = Real(c.vp) + Real(blend) * (Real(c.layer_vp) - Real(c.vp));
            const Real vs = Real(c.vs) + Real(blend) * (Real(c.layer_vs) - Real(c.vs));
            const Real rho = Real(c.density) + Real(blend) * (Real(c.layer_density) - Real(c.density));
            inv_rho[i] = Real(1) / rho;
            mu[i] = rho * vs * vs;
            lambda[i] = rho * (vp * vp - Real(2) * vs * vs);
        }
    }

    // uploadModelPlane3D (cuda/Cuda_Seis3D.cu:696-711).
    void upload_plane(int z, const float* values) {
        if (z < 0 || z >= c.nz || !values) throw Invalid("Seis3D: invalid model plane.");
        const std::size_t n = std::size_t(c.nx) * c.ny;
        for (std::size_t i = 0; i < n; ++i) {
            const float vp = values[3 * i], vs = values[3 * i + 1], rho = values[3 * i + 2];
            if (!valid_material(vp, vs, rho) || vp > c.max_vp * (1 + 1.0e-6f))
                throw Invalid("Seis3D: invalid or changed imported material.");
        }
        for (std::size_t i = 0; i < n; ++i) {
            const Real vp = values[3 * i], vs = values[3 * i + 1], rho = values[3 * i + 2];
            const int x = int(i % c.nx), y = int(i / c.nx);
            const auto j = index(x + c.ox, y + c.oy, z + c.oz);
            inv_rho[j] = Real(1) / rho; lambda[j] = rho * (vp * vp - Real(2) * vs * vs); mu[j] = rho * vs * vs;
        }
    }

    // extendMaterialsKernel3D (cuda/Cuda_Seis3D.cu:188-200) via finishModelUpload3D.
    void finish_upload() {
        for (std::size_t i = 0; i < c.cells; ++i) {
            const int x = int(i % c.gx), y = int((i / c.gx) % c.gy);
            const int z = int(i / (std::size_t(c.gx) * c.gy));
            const int px = std::max(c.ox, std::min(x, c.ox + c.nx - 1));
            const int py = std::max(c.oy, std::min(y, c.oy + c.ny - 1));
            const int pz = std::max(c.oz, std::min(z, c.oz + c.nz - 1));
            if (x == px && y == py && z == pz) continue;
            const auto j = index(px, py, pz);
            inv_rho[i] = inv_rho[j]; mu[i] = mu[j]; lambda[i] = lambda[j];
        }
    }

    // derivative3D (cuda/Cuda_Seis3D.cu:78-89).
    template <bool plus>
    Real derivative(const Real* field, std::size_t i, std::size_t stride, Real inverse_spacing) const {
        Real result = 0;
        for (int r = 0; r < 4; ++r) {
            const std::size_t positive = i + std::size_t(r + (plus ? 1 : 0)) * stride;
            const std::size_t negative = i - std::size_t(r + (plus ? 0 : 1)) * stride;
            result += Weights<Real>::w(r) * (field[positive] - field[negative]);
        }
        return result * inverse_spacing;
    }

    // applyCpml3D (cuda/Cuda_Seis3D.cu:92-111), reference slab layout.
    Real apply_cpml(int x, int y, int z, int axis, int slot, bool half, Real d) {
        if (c.reflecting) return d;
        const int n = axis == 0 ? c.gx : axis == 1 ? c.gy : c.gz;
        const int coordinate = axis == 0 ? x : axis == 1 ? y : z;
        const bool topFree = axis == 2 && c.free_surface;
        int q = -1;
        if (!topFree && coordinate < kRadius + c.pml) q = coordinate - kRadius;
        else if (coordinate >= n - kRadius - c.pml) q = (topFree ? 0 : c.pml) + coordinate - (n - kRadius - c.pml);
        if (q < 0) return d;
        const int thickness = (topFree ? 1 : 2) * c.pml;
        const std::size_t slab = axis == 0 ? (std::size_t(z) * c.gy + y) * thickness + q
            : axis == 1 ? (std::size_t(z) * c.gx + x) * thickness + q
            : (std::size_t(y) * c.gx + x) * thickness + q;
        Real& memory = cpml[axis][std::size_t(slot) * slab_cells[axis] + slab];
        const Real* k = coef[axis].data() + (std::size_t(half) * n + coordinate) * 3;
        memory = k[1] * memory + k[0] * d;
        return k[2] * d + memory;
    }

    // stressGhostKernel3D (cuda/Cuda_Seis3D.cu:203-219).
    void stress_ghost() {
        for (int y = 4; y < c.gy - 4; ++y) for (int x = 4; x < c.gx - 4; ++x) {
            f[5][index(x, y, 4)] = 0;
            for (int depth = 1; depth <= 4; ++depth) {
                const auto ghost = index(x, y, 4 - depth), mirror = index(x, y, 4 + depth);
                const auto shearMirror = index(x, y, 3 + depth);
                f[5][ghost] = -f[5][mirror];
                f[8][ghost] = -f[8][shearMirror];
                f[7][ghost] = -f[7][shearMirror];
                f[3][ghost] = f[3][mirror];
                f[4][ghost] = f[4][mirror];
                f[6][ghost] = f[6][mirror];
            }
        }
    }

    // reflectingGhostKernel3D / applyReflectingWalls3D (cuda/Cuda_Seis3D.cu:152-185).
    void reflecting_walls(bool velocities) {
        const int lengths[3] = {c.gx, c.gy, c.gz};
        for (int axis = 0; axis < 3; ++axis) {
            const int uAxis = (axis + 1) % 3, vAxis = (axis + 2) % 3;
            const int high = lengths[axis] - kRadius - 1;
            for (int v = 0; v < lengths[vAxis]; ++v) for (int u = 0; u < lengths[uAxis]; ++u) {
                int p[3]{}; p[uAxis] = u; p[vAxis] = v;
                for (int field = velocities ? 0 : 3; field < (velocities ? 3 : 9); ++field) {
                    const bool half = field < 3 ? field == axis : field == 6 ? axis != 2 : field == 7 ? axis != 0
                        : field == 8 ? axis != 1 : false;
                    const bool odd = half;
                    for (int side = 0; side < 2; ++side) for (int depth = 1; depth <= kRadius; ++depth) {
                        p[axis] = side ? high + depth : kRadius - depth;
                        const auto destination = index(p[0], p[1], p[2]);
                        if (half && !side && depth == 1) { f[field][destination] = 0; continue; }
                        p[axis] = side ? high + 1 - depth - int(half) : kRadius + depth - 1 - int(half);
                        const Real value = f[field][index(p[0], p[1], p[2])];
                        f[field][destination] = odd ? -value : value;
                    }
                    if (half) { p[axis] = high; f[field][index(p[0], p[1], p[2])] = 0; }
                }
            }
        }
    }

    // velocityKernel3D (cuda/Cuda_Seis3D.cu:222-245).
    void velocity() {
        const std::size_t strides[3] = {1, std::size_t(c.gx), std::size_t(c.gx) * c.gy};
        const int stressFields[3][3] = {{3, 6, 8}, {6, 4, 7}, {8, 7, 5}};
        const Real invH = Real(1) / Real(c.spacing), dt = Real(c.dt);
        bool bad = false;
#pragma omp parallel for schedule(static) reduction(|| : bad)
        for (int z = 4; z < c.gz - 4; ++z) for (int y = 4; y < c.gy - 4; ++y) for (int x = 4; x < c.gx - 4; ++x) {
            const auto i = index(x, y, z);
            for (int v = 0; v < 3; ++v) {
                Real divergence = 0;
                for (int axis = 0; axis < 3; ++axis) {
                    const Real* field = f[stressFields[v][axis]].data();
                    const Real d = axis == v ? derivative<true>(field, i, strides[axis], invH)
                                             : derivative<false>(field, i, strides[axis], invH);
                    divergence += apply_cpml(x, y, z, axis, v, axis == v, d);
                }
                const Real a = inv_rho[i], b = inv_rho[i + strides[v]];
                const Real value = f[v][i] + dt * (2 * a * b / (a + b)) * divergence;
                f[v][i] = value;
                if (!std::isfinite(value)) bad = true;
            }
        }
        nonfinite = nonfinite || bad;
    }

    // velocityGhostKernel3D (cuda/Cuda_Seis3D.cu:249-272).
    void velocity_ghost() {
        const Real invH = Real(1) / Real(c.spacing), h = Real(c.spacing);
        const std::size_t gx = c.gx, plane = std::size_t(c.gx) * c.gy;
        for (int y = 4; y < c.gy - 4; ++y) for (int x = 4; x < c.gx - 4; ++x) {
            const auto i = index(x, y, 4);
            const Real ex = derivative<false>(f[0].data(), i, 1, invH);
            const Real ey = derivative<false>(f[1].data(), i, gx, invH);
            const auto nextPlane = i + plane;
            const Real dzVx = -(Real(1.5) * derivative<true>(f[2].data(), i, 1, invH)
                - Real(0.5) * derivative<true>(f[2].data(), nextPlane, 1, invH));
            const Real dzVy = -(Real(1.5) * derivative<true>(f[2].data(), i, gx, invH)
                - Real(0.5) * derivative<true>(f[2].data(), nextPlane, gx, invH));
            const Real dzVz = -lambda[i] / (lambda[i] + 2 * mu[i]) * (ex + ey);
            for (int depth = 1; depth <= 4; ++depth) {
                const auto ghost = index(x, y, 4 - depth), mirror = index(x, y, 4 + depth);
                f[0][ghost] = f[0][mirror] - Real(2 * depth) * h * dzVx;
                f[1][ghost] = f[1][mirror] - Real(2 * depth) * h * dzVy;
                f[2][ghost] = f[2][index(x, y, 3 + depth)] - Real(2 * depth - 1) * h * dzVz;
            }
        }
    }

    // shearMu3D (cuda/Cuda_Seis3D.cu:275-278).
    Real shear_mu(std::size_t i, std::size_t a, std::size_t b) const {
        const Real m0 = mu[i], m1 = mu[i + a], m2 = mu[i + b], m3 = mu[i + a + b];
        return (m0 > 0 && m1 > 0 && m2 > 0 && m3 > 0) ? 4 / (1 / m0 + 1 / m1 + 1 / m2 + 1 / m3) : Real(0);
    }

    // stressKernel3D (cuda/Cuda_Seis3D.cu:281-310).
    void stress() {
        const std::size_t stride[3] = {1, std::size_t(c.gx), std::size_t(c.gx) * c.gy};
        const Real invH = Real(1) / Real(c.spacing), dt = Real(c.dt);
        bool bad = false;
#pragma omp parallel for schedule(static) reduction(|| : bad)
        for (int z = 4; z < c.gz - 4; ++z) for (int y = 4; y < c.gy - 4; ++y) for (int x = 4; x < c.gx - 4; ++x) {
            const auto i = index(x, y, z);
            Real d[3][3];
            for (int v = 0; v < 3; ++v) for (int a = 0; a < 3; ++a) {
                const Real raw = a == v ? derivative<false>(f[v].data(), i, stride[a], invH)
                                        : derivative<true>(f[v].data(), i, stride[a], invH);
                d[v][a] = apply_cpml(x, y, z, a, 3 + v, a != v, raw);
            }
            const Real lam = lambda[i], m = mu[i];
            if (c.free_surface && z == 4) d[2][2] = -lam / (lam + 2 * m) * (d[0][0] + d[1][1]);
            const Real trace = d[0][0] + d[1][1] + d[2][2];
            f[3][i] += dt * (lam * trace + 2 * m * d[0][0]);
            f[4][i] += dt * (lam * trace + 2 * m * d[1][1]);
            f[5][i] = c.free_surface && z == 4 ? Real(0) : f[5][i] + dt * (lam * trace + 2 * m * d[2][2]);
            f[6][i] += dt * shear_mu(i, stride[0], stride[1]) * (d[0][1] + d[1][0]);
            f[7][i] += dt * shear_mu(i, stride[1], stride[2]) * (d[1][2] + d[2][1]);
            f[8][i] += dt * shear_mu(i, stride[0], stride[2]) * (d[0][2] + d[2][0]);
            for (int a = 3; a < 9; ++a) if (!std::isfinite(f[a][i])) bad = true;
        }
        nonfinite = nonfinite || bad;
    }

    // Ricker value handed to the source: FP32 path rounds the FP64 series to
    // float (cuda/Cuda_Seis3D.cu:320); the FP64 oracle keeps it.
    static Real ricker(float frequency, double time) { return Real(ricker_double(frequency, time)); }

    // sourceKernel3D (cuda/Cuda_Seis3D.cu:313-332).
    void source(const seis3d_source& s, double time) {
        const Real rick = ricker(c.frequency, time);
        const Real norm = Real(10.8387785335f);
        const std::size_t strides[3] = {1, std::size_t(c.gx), std::size_t(c.gx) * c.gy};
        for (int a = 0; a < 27; ++a) {
            const int dx = a % 3 - 1, dy = (a / 3) % 3 - 1, dz = a / 9 - 1;
            const auto i = index(s.x + c.ox + dx, s.y + c.oy + dy, s.z + c.oz + dz);
            const Real w = Real(std::exp(Real(-0.5) * Real(dx * dx + dy * dy + dz * dz)));
            const Real value = Real(c.dt) * Real(s.amplitude) * rick * w / norm;
            if (s.kind == SEIS3D_EXPLOSION) {
                f[3][i] += value; f[4][i] += value; f[5][i] += value;
            } else {
                for (int v = 0; v < 3; ++v) {
                    const Real a_ = inv_rho[i], b_ = inv_rho[i + strides[v]];
                    f[v][i] += value * Real(s.direction[v]) * (2 * a_ * b_ / (a_ + b_));
                }
            }
        }
    }

    // runStep3D argument checks (cuda/Cuda_Seis3D.cu:603-607).
    void check_source(const seis3d_source& s) const {
        if (s.x < 1 || s.x >= c.nx - 1 || s.y < 1 || s.y >= c.ny - 1 || s.z < 2 || s.z >= c.nz - 1
            || !std::isfinite(s.amplitude))
            throw Invalid("Seis3D: source must be inside the physical volume, at depth >=2.");
        for (float d : s.direction) if (!std::isfinite(d)) throw Invalid("Seis3D: invalid force direction.");
    }

    // runStep3D (cuda/Cuda_Seis3D.cu:602-624): the launch order of one step.
    void step(const seis3d_source& s, std::uint64_t n) {
        check_source(s);
        if (c.free_surface) stress_ghost();
        if (c.reflecting) reflecting_walls(false);
        velocity();
        if (s.kind == SEIS3D_FORCE) source(s, double(n) * c.dt);
        if (c.free_surface) velocity_ghost();
        if (c.reflecting) reflecting_walls(true);
        stress();
        if (s.kind == SEIS3D_EXPLOSION) source(s, (double(n) + 0.5) * c.dt);
        if (c.free_surface) stress_ghost();
        if (c.reflecting) reflecting_walls(false);
    }

    // setReceivers3D validation (cuda/Cuda_Seis3D.cu:649-655).
    void set_receivers(const seis3d_receiver* r, int count) {
        if (count < 0 || count > kMaxReceivers || (count && !r)) throw Invalid("Seis3D: invalid receiver count.");
        for (int a = 0; a < count; ++a)
            if (r[a].x < 0 || r[a].x >= c.nx || r[a].y < 0 || r[a].y >= c.ny || r[a].z < 0 || r[a].z >= c.nz)
                throw Invalid("Seis3D: receiver outside the physical domain.");
        receivers.assign(r, r + count);
    }

    // sampleScalar3D (cuda/Cuda_Seis3D.cu:335-352) at a physical coordinate.
    Real sample_scalar(int x, int y, int z, int component) const {
        const auto i = index(x + c.ox, y + c.oy, z + c.oz);
        const std::size_t gx = c.gx, plane = std::size_t(c.gx) * c.gy;
        if (component == 5) return std::sqrt((lambda[i] + 2 * mu[i]) * inv_rho[i]);
        if (component == 6) return std::sqrt(mu[i] * inv_rho[i]);
        if (component == 7) return 1 / inv_rho[i];
        const Real vx = Real(0.5) * (f[0][i] + f[0][i - 1]);
        const Real vy = Real(0.5) * (f[1][i] + f[1][i - gx]);
        const Real vz = Real(0.5) * (f[2][i] + f[2][i - plane]);
        const Real magnitude = std::sqrt(vx * vx + vy * vy + vz * vz);
        if (component == 0) return vz;
        if (component == 1) return vx;
        if (component == 2) return vy;
        if (component == 3) return magnitude;
        if (component == 4) return -(f[3][i] + f[4][i] + f[5][i]) / 3;
        return 0;
    }

    // receiverKernel3D (cuda/Cuda_Seis3D.cu:421-428): [r][vx, vy, vz].
    void sample(Real* out) const {
        for (std::size_t r = 0; r < receivers.size(); ++r) {
            const auto p = receivers[r];
            out[3 * r] = sample_scalar(p.x, p.y, p.z, 1);
            out[3 * r + 1] = sample_scalar(p.x, p.y, p.z, 2);
            out[3 * r + 2] = sample_scalar(p.x, p.y, p.z, 0);
        }
    }

    // ------------------------------------------------------------------
    // Discrete adjoint of one step (no reference counterpart).
    //
    // The step is linear in (v, sigma, psi) for a fixed model, so its adjoint
    // is the transpose of the same operations in reverse order: the stencil
    // transposes are D+^T = -D- and D-^T = -D+, the CPML recurrence
    // psi' = b psi + a d, e = d/kappa + psi' transposes to
    // L_tot = L_psi' + L_e, L_d = L_e/kappa + a L_tot, L_psi = b L_tot,
    // and the constitutive law to L_e = dt M^T L_sigma.  Model gradients are
    // taken with respect to 1/rho, lambda and mu at every padded cell,
    // including the harmonic face buoyancy 2ab/(a+b) and the 4-cell harmonic
    // shear modulus.  The free surface's ghost fills and the reflecting walls are
    // linear maps too: their transposes move the ghost adjoints back onto the
    // cells the ghosts were built from, the velocity ghosts of a free surface
    // adding a lambda/(lambda+2mu) gradient term.
    // ------------------------------------------------------------------
    struct Snapshot {
        std::vector<Real> f[9];
        std::vector<Real> cpml[3];
    };
    Snapshot save() const {
        Snapshot s;
        for (int k = 0; k < 9; ++k) s.f[k] = f[k];
        for (int a = 0; a < 3; ++a) s.cpml[a] = cpml[a];
        return s;
    }
    void load(const Snapshot& s) {
        for (int k = 0; k < 9; ++k) f[k] = s.f[k];
        for (int a = 0; a < 3; ++a) cpml[a] = s.cpml[a];
    }
    struct Adjoint {
        std::vector<Real> f[9];      // adjoints of the nine wavefields
        std::vector<Real> cpml[3];   // adjoints of the CPML memories (same slab layout)
        std::vector<Real> g_ib, g_lam;    // dJ/d(1/rho), dJ/dlambda (= dJ/dM at fixed mu, M = lambda + 2 mu)
        std::vector<Real> g_dev;          // dJ/dmu at fixed M = dJ/dmu - 2 dJ/dlambda
    };
    Adjoint make_adjoint() const {
        Adjoint A;
        for (auto& v : A.f) v.assign(c.cells, Real(0));
        for (int a = 0; a < 3; ++a) A.cpml[a].assign(cpml[a].size(), Real(0));
        A.g_ib.assign(c.cells, Real(0)); A.g_lam.assign(c.cells, Real(0)); A.g_dev.assign(c.cells, Real(0));
        return A;
    }

    // Slab position of a derivative, -1 when outside the axis' CPML slab.
    long long slab_slot_index(int x, int y, int z, int axis, int slot, int& coordinate) const {
        const int n = axis == 0 ? c.gx : axis == 1 ? c.gy : c.gz;
        coordinate = axis == 0 ? x : axis == 1 ? y : z;
        const bool topFree = axis == 2 && c.free_surface;
        int q = -1;
        if (!topFree && coordinate < kRadius + c.pml) q = coordinate - kRadius;
        else if (coordinate >= n - kRadius - c.pml) q = (topFree ? 0 : c.pml) + coordinate - (n - kRadius - c.pml);
        if (q < 0 || c.reflecting) return -1;
        const int thickness = (topFree ? 1 : 2) * c.pml;
        const std::size_t slab = axis == 0 ? (std::size_t(z) * c.gy + y) * thickness + q
            : axis == 1 ? (std::size_t(z) * c.gx + x) * thickness + q
            : (std::size_t(y) * c.gx + x) * thickness + q;
        return (long long)(std::size_t(slot) * slab_cells[axis] + slab);
    }

    // Filtered derivative e = d/kappa + psi' of a stored state (psi' as stored).
    Real filtered(const Snapshot& after, Real d, int x, int y, int z, int axis, int slot, bool half) const {
        int coordinate = 0;
        const long long k = slab_slot_index(x, y, z, axis, slot, coordinate);
        if (k < 0) return d;
        const int n = axis == 0 ? c.gx : axis == 1 ? c.gy : c.gz;
        const Real* kk = coef[axis].data() + (std::size_t(half) * n + coordinate) * 3;
        return kk[2] * d + after.cpml[axis][std::size_t(k)];
    }

    // Adjoint of the CPML filter for one derivative: returns L_d and updates L_psi.
    Real adjoint_filter(Adjoint& A, Real le, int x, int y, int z, int axis, int slot, bool half) const {
        int coordinate = 0;
        const long long k = slab_slot_index(x, y, z, axis, slot, coordinate);
        if (k < 0) return le;
        const int n = axis == 0 ? c.gx : axis == 1 ? c.gy : c.gz;
        const Real* kk = coef[axis].data() + (std::size_t(half) * n + coordinate) * 3;
        Real& lp = A.cpml[axis][std::size_t(k)];
        const Real total = lp + le;
        lp = kk[1] * total;
        return kk[2] * le + kk[0] * total;
    }

    // Stencil along an axis at a ghost cell (coordinate `coord` of `n`) of a temporary
    // that is zero outside the updated cells: taps outside [4, n - 4) are zero, and
    // are not read (they may lie outside the grid).
    template <bool plus>
    Real clipped_derivative(const Real* field, std::size_t i, int coord, int n, std::size_t stride,
                            Real inverse_spacing) const {
        Real result = 0;
        for (int r = 0; r < 4; ++r) {
            const int up = r + (plus ? 1 : 0), down = r + (plus ? 0 : 1);
            const Real fp = coord + up >= 4 && coord + up < n - 4 ? field[i + std::size_t(up) * stride] : Real(0);
            const Real fn = coord - down >= 4 && coord - down < n - 4 ? field[i - std::size_t(down) * stride] : Real(0);
            result += Weights<Real>::w(r) * (fp - fn);
        }
        return result * inverse_spacing;
    }

    // The stencils of the cells next to a boundary read its ghost cells: the four
    // planes above a free surface, or every face's with reflecting walls.  The
    // transposed stencils gather the temporaries L_d into those ghost adjoints (only
    // the stencil along the face normal reaches them): L_v after the stress half-step,
    // L_sigma after the velocity half-step.
    void ghost_gathers(const std::vector<Real> (&ld)[9], Adjoint& A, bool velocity_pass) const {
        const int stressFields[3][3] = {{3, 6, 8}, {6, 4, 7}, {8, 7, 5}};
        const int n[3] = {c.gx, c.gy, c.gz};
        const std::size_t st[3] = {1, std::size_t(c.gx), std::size_t(c.gx) * c.gy};
        const Real invH = Real(1) / Real(c.spacing);
        for (int axis = 0; axis < 3; ++axis) for (int side = 0; side < 2; ++side) {
            if (!c.reflecting && !(c.free_surface && axis == 2 && side == 0)) continue;
            int lo[3], hi[3];
            for (int a = 0; a < 3; ++a) { lo[a] = 4; hi[a] = n[a] - 4; }
            lo[axis] = side ? n[axis] - 4 : 0;
            hi[axis] = side ? n[axis] : 4;
            for (int z = lo[2]; z < hi[2]; ++z) for (int y = lo[1]; y < hi[1]; ++y) for (int x = lo[0]; x < hi[0]; ++x) {
                const auto i = index(x, y, z);
                const int coord = axis == 0 ? x : axis == 1 ? y : z;
                for (int v = 0; v < 3; ++v) {
                    const bool plus = velocity_pass ? axis != v : axis == v;
                    const Real* t = ld[3 * v + axis].data();
                    const Real d = plus ? clipped_derivative<true>(t, i, coord, n[axis], st[axis], invH)
                                        : clipped_derivative<false>(t, i, coord, n[axis], st[axis], invH);
                    A.f[velocity_pass ? stressFields[v][axis] : v][i] -= d;
                }
            }
        }
    }

    // Transpose of reflecting_walls: axes in reverse order; on each line the zeroed
    // cells' adjoints clear and each ghost adjoint moves onto its source (with the
    // ghost's sign) and clears.
    void adjoint_reflecting_walls(Adjoint& A, bool velocities) const {
        const int lengths[3] = {c.gx, c.gy, c.gz};
        for (int axis = 2; axis >= 0; --axis) {
            const int uAxis = (axis + 1) % 3, vAxis = (axis + 2) % 3;
            const int high = lengths[axis] - kRadius - 1;
            for (int v = 0; v < lengths[vAxis]; ++v) for (int u = 0; u < lengths[uAxis]; ++u) {
                int p[3]{}; p[uAxis] = u; p[vAxis] = v;
                for (int field = velocities ? 0 : 3; field < (velocities ? 3 : 9); ++field) {
                    const bool half = field < 3 ? field == axis : field == 6 ? axis != 2 : field == 7 ? axis != 0
                        : field == 8 ? axis != 1 : false;
                    Real* L = A.f[field].data();
                    if (half) { p[axis] = high; L[index(p[0], p[1], p[2])] = 0; }
                    for (int side = 1; side >= 0; --side) for (int depth = kRadius; depth >= 1; --depth) {
                        p[axis] = side ? high + depth : kRadius - depth;
                        const auto destination = index(p[0], p[1], p[2]);
                        if (!(half && !side && depth == 1)) {
                            p[axis] = side ? high + 1 - depth - int(half) : kRadius + depth - 1 - int(half);
                            const auto source = index(p[0], p[1], p[2]);
                            L[source] += half ? -L[destination] : L[destination];
                        }
                        L[destination] = 0;
                    }
                }
            }
        }
    }

    // Adds c * d(derivative<plus>(field, i, stride))/d(field) to the adjoint `out`.
    template <bool plus>
    void scatter_derivative(Real* out, std::size_t i, std::size_t stride, Real c_) const {
        for (int r = 0; r < 4; ++r) {
            out[i + std::size_t(r + (plus ? 1 : 0)) * stride] += c_ * Weights<Real>::w(r);
            out[i - std::size_t(r + (plus ? 0 : 1)) * stride] -= c_ * Weights<Real>::w(r);
        }
    }

    // Transpose of stress_ghost: each ghost adjoint moves onto the cell it mirrors
    // (with the ghost's sign) and is cleared; then the surface szz adjoint is cleared.
    void adjoint_stress_ghost(Adjoint& A) const {
        for (int y = 4; y < c.gy - 4; ++y) for (int x = 4; x < c.gx - 4; ++x) {
            for (int depth = 1; depth <= 4; ++depth) {
                const auto ghost = index(x, y, 4 - depth), mirror = index(x, y, 4 + depth);
                const auto shearMirror = index(x, y, 3 + depth);
                A.f[5][mirror] -= A.f[5][ghost];
                A.f[8][shearMirror] -= A.f[8][ghost];
                A.f[7][shearMirror] -= A.f[7][ghost];
                A.f[3][mirror] += A.f[3][ghost];
                A.f[4][mirror] += A.f[4][ghost];
                A.f[6][mirror] += A.f[6][ghost];
                for (int k : {3, 4, 5, 6, 7, 8}) A.f[k][ghost] = 0;
            }
            A.f[5][index(x, y, 4)] = 0;
        }
    }

    // Transpose of velocity_ghost on the velocities `v` the ghosts were built from.
    // The ghosts are linear in the velocities and, through r = lambda/(lambda+2mu),
    // in the model: dJ/dr = -(ex + ey) L_dzVz at each surface cell.
    void adjoint_velocity_ghost(const Snapshot& v, Adjoint& A) const {
        const Real invH = Real(1) / Real(c.spacing), h = Real(c.spacing);
        const std::size_t gx = c.gx, plane = std::size_t(c.gx) * c.gy;
        for (int y = 4; y < c.gy - 4; ++y) for (int x = 4; x < c.gx - 4; ++x) {
            const auto i = index(x, y, 4);
            Real l_dzvx = 0, l_dzvy = 0, l_dzvz = 0;
            for (int depth = 1; depth <= 4; ++depth) {
                const auto ghost = index(x, y, 4 - depth), mirror = index(x, y, 4 + depth);
                const Real g0 = A.f[0][ghost], g1 = A.f[1][ghost], g2 = A.f[2][ghost];
                A.f[0][mirror] += g0;
                A.f[1][mirror] += g1;
                A.f[2][index(x, y, 3 + depth)] += g2;
                l_dzvx -= Real(2 * depth) * h * g0;
                l_dzvy -= Real(2 * depth) * h * g1;
                l_dzvz -= Real(2 * depth - 1) * h * g2;
                A.f[0][ghost] = 0; A.f[1][ghost] = 0; A.f[2][ghost] = 0;
            }
            // dzVx = -1.5 D+x vz(z=4) + 0.5 D+x vz(z=5), dzVy likewise along y.
            scatter_derivative<true>(A.f[2].data(), i, 1, Real(-1.5) * invH * l_dzvx);
            scatter_derivative<true>(A.f[2].data(), i + plane, 1, Real(0.5) * invH * l_dzvx);
            scatter_derivative<true>(A.f[2].data(), i, gx, Real(-1.5) * invH * l_dzvy);
            scatter_derivative<true>(A.f[2].data(), i + plane, gx, Real(0.5) * invH * l_dzvy);
            // dzVz = -r (ex + ey), ex = D-x vx, ey = D-y vy at the surface.
            const Real lam = lambda[i], m = mu[i], den = lam + 2 * m;
            const Real r = lam / den;
            scatter_derivative<false>(A.f[0].data(), i, 1, -r * invH * l_dzvz);
            scatter_derivative<false>(A.f[1].data(), i, gx, -r * invH * l_dzvz);
            const Real ex = derivative<false>(v.f[0].data(), i, 1, invH);
            const Real ey = derivative<false>(v.f[1].data(), i, gx, invH);
            const Real gr = -(ex + ey) * l_dzvz;
            A.g_lam[i] += gr * 2 * m / (den * den);
            A.g_dev[i] -= gr * 2 / den;                     // dr/dmu at fixed lambda + 2 mu
        }
    }

    // Adjoint of one full step n (state `before` -> `after`), with the misfit
    // residual of a sample taken after this step already added to A.f[0..2].
    void adjoint_step(const Snapshot& before, const Snapshot& after, Adjoint& A, const seis3d_source& s,
                      std::uint64_t n) {
        const std::size_t st[3] = {1, std::size_t(c.gx), std::size_t(c.gx) * c.gy};
        const Real invH = Real(1) / Real(c.spacing), dt = Real(c.dt);
        const bool free = c.free_surface != 0, walls = c.reflecting != 0;
        std::vector<Real> ld[9];
        for (auto& v : ld) v.assign(c.cells, Real(0));
        if (free) adjoint_stress_ghost(A);
        if (walls) adjoint_reflecting_walls(A, false);
        // ---- stress half-step ----
        for (int z = 4; z < c.gz - 4; ++z) for (int y = 4; y < c.gy - 4; ++y) for (int x = 4; x < c.gx - 4; ++x) {
            const auto i = index(x, y, z);
            Real e[3][3];
            for (int v = 0; v < 3; ++v) for (int a = 0; a < 3; ++a) {
                const Real d = a == v ? derivative<false>(after.f[v].data(), i, st[a], invH)
                                      : derivative<true>(after.f[v].data(), i, st[a], invH);
                e[v][a] = filtered(after, d, x, y, z, a, 3 + v, a != v);
            }
            const Real lam = lambda[i], m = mu[i];
            const Real Ls[6] = {A.f[3][i], A.f[4][i], A.f[5][i], A.f[6][i], A.f[7][i], A.f[8][i]};
            Real le[3][3] = {};
            if (free && z == 4) {
                // Surface law: d_zz = -r (e_xx + e_yy) makes lambda trace = l' (e_xx + e_yy) with
                // l' = 2 lambda mu / (lambda + 2 mu); szz is set to 0, so its adjoint clears.
                const Real den = lam + 2 * m, lp = 2 * lam * m / den;
                const Real sumL = Ls[0] + Ls[1], exy = e[0][0] + e[1][1];
                le[0][0] = dt * (lp * sumL + 2 * m * Ls[0]);
                le[1][1] = dt * (lp * sumL + 2 * m * Ls[1]);
                const Real glp = dt * exy * sumL;
                A.g_lam[i] += glp * 4 * m * m / (den * den);
                A.g_dev[i] += glp * 2 * (lam - 2 * m) / den + dt * 2 * (e[0][0] * Ls[0] + e[1][1] * Ls[1]);
                A.f[5][i] = 0;
            } else {
                const Real sumL = Ls[0] + Ls[1] + Ls[2];
                for (int v = 0; v < 3; ++v) le[v][v] = dt * (lam * sumL + 2 * m * Ls[v]);
                const Real tr = e[0][0] + e[1][1] + e[2][2];
                A.g_lam[i] += dt * tr * sumL;
                A.g_dev[i] -= dt * 2 * (e[0][0] * (Ls[1] + Ls[2]) + e[1][1] * (Ls[0] + Ls[2]) + e[2][2] * (Ls[0] + Ls[1]));
            }
            const Real mxy = shear_mu(i, st[0], st[1]), myz = shear_mu(i, st[1], st[2]), mxz = shear_mu(i, st[0], st[2]);
            le[0][1] = le[1][0] = dt * mxy * Ls[3];
            le[1][2] = le[2][1] = dt * myz * Ls[4];
            le[0][2] = le[2][0] = dt * mxz * Ls[5];
            const Real gh[3] = {dt * (e[0][1] + e[1][0]) * Ls[3], dt * (e[1][2] + e[2][1]) * Ls[4],
                                dt * (e[0][2] + e[2][0]) * Ls[5]};
            const std::size_t pairs[3][2] = {{st[0], st[1]}, {st[1], st[2]}, {st[0], st[2]}};
            const Real mh[3] = {mxy, myz, mxz};
            for (int p = 0; p < 3; ++p) {
                if (mh[p] == 0) continue;
                const std::size_t cells4[4] = {i, i + pairs[p][0], i + pairs[p][1], i + pairs[p][0] + pairs[p][1]};
                for (auto k : cells4) A.g_dev[k] += gh[p] * mh[p] * mh[p] / (4 * mu[k] * mu[k]);
            }
            for (int v = 0; v < 3; ++v) for (int a = 0; a < 3; ++a)
                ld[3 * v + a][i] = adjoint_filter(A, le[v][a], x, y, z, a, 3 + v, a != v);
        }
        // L_v += D^T L_d (transpose of D- for a == v, of D+ otherwise).
        for (int z = 4; z < c.gz - 4; ++z) for (int y = 4; y < c.gy - 4; ++y) for (int x = 4; x < c.gx - 4; ++x) {
            const auto i = index(x, y, z);
            for (int v = 0; v < 3; ++v) for (int a = 0; a < 3; ++a) {
                const Real t = a == v ? derivative<true>(ld[3 * v + a].data(), i, st[a], invH)
                                      : derivative<false>(ld[3 * v + a].data(), i, st[a], invH);
                A.f[v][i] -= t;
            }
        }
        ghost_gathers(ld, A, false);
        for (auto& v : ld) std::fill(v.begin(), v.end(), Real(0));
        if (free) adjoint_velocity_ghost(after, A);
        if (walls) adjoint_reflecting_walls(A, true);
        // ---- velocity half-step ----
        const int stressFields[3][3] = {{3, 6, 8}, {6, 4, 7}, {8, 7, 5}};
        const Real rick = ricker(c.frequency, double(n) * c.dt);
        const Real norm = Real(10.8387785335f);
        for (int z = 4; z < c.gz - 4; ++z) for (int y = 4; y < c.gy - 4; ++y) for (int x = 4; x < c.gx - 4; ++x) {
            const auto i = index(x, y, z);
            for (int v = 0; v < 3; ++v) {
                Real E = 0;
                for (int a = 0; a < 3; ++a) {
                    const Real* fld = before.f[stressFields[v][a]].data();
                    const Real d = a == v ? derivative<true>(fld, i, st[a], invH) : derivative<false>(fld, i, st[a], invH);
                    E += filtered(after, d, x, y, z, a, v, a == v);
                }
                const Real lv = A.f[v][i];
                const Real aa = inv_rho[i], bb = inv_rho[i + st[v]];
                const Real h = 2 * aa * bb / (aa + bb);
                Real dvdh = dt * E;
                if (s.kind == SEIS3D_FORCE) {
                    const int dx = x - (s.x + c.ox), dy = y - (s.y + c.oy), dz = z - (s.z + c.oz);
                    if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1) {
                        const Real w = Real(std::exp(Real(-0.5) * Real(dx * dx + dy * dy + dz * dz)));
                        dvdh += Real(c.dt) * Real(s.amplitude) * rick * w / norm * Real(s.direction[v]);
                    }
                }
                const Real gh = dvdh * lv;
                const Real den = (aa + bb) * (aa + bb);
                A.g_ib[i] += gh * 2 * bb * bb / den;
                A.g_ib[i + st[v]] += gh * 2 * aa * aa / den;
                const Real le = dt * h * lv;
                for (int a = 0; a < 3; ++a) ld[3 * v + a][i] = adjoint_filter(A, le, x, y, z, a, v, a == v);
            }
        }
        // L_sigma += D^T L_d (transpose of D+ for a == v, of D- otherwise).
        for (int z = 4; z < c.gz - 4; ++z) for (int y = 4; y < c.gy - 4; ++y) for (int x = 4; x < c.gx - 4; ++x) {
            const auto i = index(x, y, z);
            for (int v = 0; v < 3; ++v) for (int a = 0; a < 3; ++a) {
                const Real t = a == v ? derivative<false>(ld[3 * v + a].data(), i, st[a], invH)
                                      : derivative<true>(ld[3 * v + a].data(), i, st[a], invH);
                A.f[stressFields[v][a]][i] -= t;
            }
        }
        ghost_gathers(ld, A, true);
        // The step opens with stress ghosts and walls too (no-ops on a stored state, not
        // on an arbitrary one).
        if (free) adjoint_stress_ghost(A);
        if (walls) adjoint_reflecting_walls(A, false);
    }

    // Residual injection R^T r for a sample (collocation adjoint).
    void inject_residual(Adjoint& A, const Real* r) const {
        const std::size_t st[3] = {1, std::size_t(c.gx), std::size_t(c.gx) * c.gy};
        const int comp[3] = {0, 1, 2};   // r is [vx, vy, vz]
        for (std::size_t k = 0; k < receivers.size(); ++k) {
            const auto p = receivers[k];
            const auto i = index(p.x + c.ox, p.y + c.oy, p.z + c.oz);
            for (int v = 0; v < 3; ++v) {
                const Real h = Real(0.5) * r[3 * k + comp[v]];
                A.f[v][i] += h;
                A.f[v][i - st[v]] += h;
            }
        }
    }

    // Misfit 1/2 sum (gather - observed)^2 and its gradient with respect to
    // vp, vs, rho at every physical cell ([z][y][x]), storing every state.
    Real fwi_gradient(const seis3d_source& s, std::int64_t steps, int stride, const Real* observed,
                      Real* g_vp, Real* g_vs, Real* g_rho) {
        reset();
        std::vector<Snapshot> states;
        states.reserve(std::size_t(steps) + 1);
        states.push_back(save());
        const std::size_t nr = receivers.size();
        std::vector<Real> gather(std::size_t(steps / stride) * nr * 3);
        int count = 0;
        for (std::int64_t n = 0; n < steps; ++n) {
            step(s, std::uint64_t(n));
            states.push_back(save());
            if ((n + 1) % stride == 0) sample(gather.data() + std::size_t(count++) * nr * 3);
        }
        Real misfit = 0;
        std::vector<Real> res(gather.size());
        for (std::size_t k = 0; k < gather.size(); ++k) {
            res[k] = gather[k] - observed[k];
            misfit += Real(0.5) * res[k] * res[k];
        }
        Adjoint A = make_adjoint();
        for (std::int64_t n = steps - 1; n >= 0; --n) {
            if ((n + 1) % stride == 0) inject_residual(A, res.data() + std::size_t((n + 1) / stride - 1) * nr * 3);
            adjoint_step(states[std::size_t(n)], states[std::size_t(n) + 1], A, s, std::uint64_t(n));
        }
        load(states.back());
        // Fold the padding (materials are clamp-extended copies) onto physical cells.
        std::vector<Real> gib(std::size_t(c.nx) * c.ny * c.nz, 0), glam(gib.size(), 0), gdev(gib.size(), 0);
        for (std::size_t i = 0; i < c.cells; ++i) {
            const int x = int(i % c.gx), y = int((i / c.gx) % c.gy), z = int(i / (std::size_t(c.gx) * c.gy));
            const int px = std::max(0, std::min(c.nx - 1, x - c.ox));
            const int py = std::max(0, std::min(c.ny - 1, y - c.oy));
            const int pz = std::max(0, std::min(c.nz - 1, z - c.oz));
            const std::size_t j = (std::size_t(pz) * c.ny + py) * c.nx + px;
            gib[j] += A.g_ib[i]; glam[j] += A.g_lam[i]; gdev[j] += A.g_dev[i];
        }
        for (int z = 0; z < c.nz; ++z) for (int y = 0; y < c.ny; ++y) for (int x = 0; x < c.nx; ++x) {
            const std::size_t j = (std::size_t(z) * c.ny + y) * c.nx + x;
            const auto i = index(x + c.ox, y + c.oy, z + c.oz);
            const Real rho = Real(1) / inv_rho[i];
            const Real vs2 = mu[i] / rho, vp2 = (lambda[i] + 2 * mu[i]) / rho;
            const Real vp = std::sqrt(vp2), vs = std::sqrt(vs2);
            g_vp[j] = glam[j] * 2 * rho * vp;
            // M = lambda + 2 mu = rho vp^2 (g_lam = dJ/dM at fixed mu), mu = rho vs^2 at fixed M (g_dev).
            g_vs[j] = gdev[j] * 2 * rho * vs;
            g_rho[j] = glam[j] * vp2 + gdev[j] * vs2 - gib[j] / (rho * rho);
        }
        return misfit;
    }
};

// fitSurvey3D (src/Seis3D_Sim.cpp:42-60).
inline void fit_survey(const seis3d_ctx& c, seis3d_survey& s) {
    s.receiver_count_x = std::clamp(s.receiver_count_x, 1, std::min(kMaxReceivers, c.nx - 2));
    s.receiver_count_y = std::clamp(s.receiver_count_y, 1, std::min(kMaxReceivers / s.receiver_count_x, c.ny - 2));
    s.shot_count_x = std::clamp(s.shot_count_x, 1, 100);
    s.shot_count_y = std::clamp(s.shot_count_y, 1, 100);
    s.receiver_origin[2] = 0;
    s.receiver_spacing[0] = std::max(1, (c.nx - 3) / std::max(1, s.receiver_count_x - 1));
    s.receiver_spacing[1] = std::max(1, (c.ny - 3) / std::max(1, s.receiver_count_y - 1));
    s.receiver_origin[0] = s.receiver_count_x == 1 ? c.nx / 2
        : (c.nx - 1 - (s.receiver_count_x - 1) * s.receiver_spacing[0]) / 2;
    s.receiver_origin[1] = s.receiver_count_y == 1 ? c.ny / 2
        : (c.ny - 1 - (s.receiver_count_y - 1) * s.receiver_spacing[1]) / 2;
    s.shot_origin[0] = c.nx / 4; s.shot_origin[1] = c.ny / 4; s.shot_origin[2] = std::max(2, c.nz / 8);
    s.shot_spacing[0] = std::max(1, (c.nx - 2 - s.shot_origin[0]) / std::max(1, s.shot_count_x - 1));
    s.shot_spacing[1] = std::max(1, (c.ny - 2 - s.shot_origin[1]) / std::max(1, s.shot_count_y - 1));
}

// configureReceivers3D (src/Seis3D/Acquisition3D.cpp:21-37): y-major, then x.
inline std::vector<seis3d_receiver> survey_receivers(const seis3d_ctx& c, const seis3d_survey& q) {
    if (q.receiver_count_x < 1 || q.receiver_count_y < 1 || q.receiver_count_x > kMaxReceivers
        || q.receiver_count_y > kMaxReceivers || std::int64_t(q.receiver_count_x) * q.receiver_count_y > kMaxReceivers
        || q.receiver_spacing[0] < 1 || q.receiver_spacing[1] < 1)
        throw std::runtime_error("Receiver mesh must have 1..100 total channels (X times Y) and positive spacing.");
    const auto lastX = std::int64_t(q.receiver_origin[0]) + std::int64_t(q.receiver_count_x - 1) * q.receiver_spacing[0];
    const auto lastY = std::int64_t(q.receiver_origin[1]) + std::int64_t(q.receiver_count_y - 1) * q.receiver_spacing[1];
    if (q.receiver_origin[0] < 0 || q.receiver_origin[1] < 0 || lastX >= c.nx || lastY >= c.ny
        || q.receiver_origin[2] < 0 || q.receiver_origin[2] >= c.nz)
        throw std::runtime_error("Receiver mesh exceeds the model.");
    std::vector<seis3d_receiver> out;
    for (int y = 0; y < q.receiver_count_y; ++y) for (int x = 0; x < q.receiver_count_x; ++x)
        out.push_back({q.receiver_origin[0] + x * q.receiver_spacing[0], q.receiver_origin[1] + y * q.receiver_spacing[1],
                       q.receiver_origin[2]});
    return out;
}

}  // namespace seis3d_oracle
