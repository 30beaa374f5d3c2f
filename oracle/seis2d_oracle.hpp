// seis2d_oracle.hpp - CPU restatement of the reference's 2D P-SV solver
// (CUDA_Kernels/Cuda_Check.cu:225-667,793-834,841-1085), templated on the
// arithmetic type (float = the reference FP32 path with IEEE arithmetic,
// double = the scheme).  Expressions keep the reference's operand order so
// that, built without FP contraction, the FP32 restatement is the exact IEEE
// evaluation of each kernel.
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP 2D solver and the timed
// CPU baseline of bench.py; never a product path.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "../include/seis2d.h"

namespace seis2d_oracle {

constexpr int kRadius = 4;                   // Main_header/SolverNumerics.h:9
constexpr int kMaxSources = 200;             // SolverNumerics.h:10
constexpr float kSourceScale = 1e7f;         // SolverNumerics.h:12

struct Invalid : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

// Host arrays of one SimulationContext (Main_header/Common.h:259-287).
struct Context {
    seis2d_ctx c{};
    std::vector<float> rho, mu, lambda, lambda2mu, dp_flat, dx, dx_half, dz, dz_half, wavelet;
    std::vector<int> rcv;
};

inline Context context_from(const seis2d_ctx& c, const seis2d_arrays& a) {
    Context o;
    o.c = c;
    const std::size_t n = std::size_t(c.nx) * c.nz;
    auto take = [](const float* p, std::size_t k) { return p ? std::vector<float>(p, p + k) : std::vector<float>(); };
    o.rho = take(a.rho, n); o.mu = take(a.mu, n); o.lambda = take(a.lambda, n); o.lambda2mu = take(a.lambda2mu, n);
    o.dp_flat = take(a.dp_flat, c.flag_type == 3 ? std::size_t(c.nx) : 0);
    o.dx = take(a.dx, c.nx); o.dx_half = take(a.dx_half, c.nx); o.dz = take(a.dz, c.nz); o.dz_half = take(a.dz_half, c.nz);
    o.wavelet = take(a.wavelet, std::size_t(std::max(0, a.wavelet_len)));
    if (a.rcv_grid_idx && c.num_rcv > 0) o.rcv.assign(a.rcv_grid_idx, a.rcv_grid_idx + c.num_rcv);
    return o;
}

// checkedGridElementCount / validateBoundaryGeometry (SolverNumerics.h:25-46).
inline std::size_t grid_count(int nx, int nz) {
    if (nx <= 2 * kRadius || nz <= 2 * kRadius) throw Invalid("Grid dimensions must exceed the 8-cell stencil guard region");
    const std::size_t n = std::size_t(nx) * std::size_t(nz);
    if (n > std::size_t(2147483647)) throw Invalid("The 2D solver index capacity was exceeded");
    return n;
}
inline void boundary_geometry(int nx, int nz, int pml, bool free_surface) {
    grid_count(nx, nz);
    if (pml < 0 || pml > (nx - 2 * kRadius - 1) / 2 || pml > (nz - 2 * kRadius - 1) / 2)
        throw Invalid("PML width leaves no physical interior grid");
    if (free_surface && pml < kRadius) throw Invalid("A free surface requires at least 4 top ghost rows");
}

inline float courant_limit() {
    const double s = 1225.0 / 1024 + 245.0 / 3072 + 49.0 / 5120 + 5.0 / 7168;
    return float(1.0 / (1.4142135623730951 * s));
}

// validateSimulationContext (Cuda_Check.cu:841-884).
inline void validate(const Context& x) {
    const seis2d_ctx& c = x.c;
    const std::size_t n = grid_count(c.nx, c.nz);
    boundary_geometry(c.nx, c.nz, c.npml, c.flag_type == 3 && c.up_flag);
    if (c.flag_type < 1 || c.flag_type > 3 || std::size_t(c.total_grid) != n || !std::isfinite(c.dt) || c.dt <= 0 ||
        !std::isfinite(c.h) || c.h <= 0 || c.nt <= 0)
        throw Invalid("Invalid solver, time grid, or spatial grid");
    if (x.rho.size() != n || x.mu.size() != n || x.lambda.size() != n || x.lambda2mu.size() != n ||
        x.dx.size() != std::size_t(c.nx) || x.dx_half.size() != std::size_t(c.nx) || x.dz.size() != std::size_t(c.nz) ||
        x.dz_half.size() != std::size_t(c.nz) || (c.flag_type == 3 && x.dp_flat.size() != std::size_t(c.nx)) ||
        x.wavelet.size() < std::size_t(c.nt) || c.num_rcv < 0 || x.rcv.size() != std::size_t(c.num_rcv))
        throw Invalid("Simulation array lengths do not match the grid or receiver count");
    float vmax = 0;
    for (std::size_t k = 0; k < n; ++k) {
        if (!std::isfinite(x.rho[k]) || x.rho[k] <= 0 || !std::isfinite(x.mu[k]) || x.mu[k] < 0 ||
            !std::isfinite(x.lambda[k]) || !std::isfinite(x.lambda2mu[k]) || x.lambda2mu[k] <= 0)
            throw Invalid("Density/stiffness must be finite with rho > 0, mu >= 0, lambda2mu > 0");
        vmax = std::max(vmax, std::sqrt(x.lambda2mu[k] / x.rho[k]));
        if (x.lambda2mu[k] - (4.0f / 3.0f) * x.mu[k] <= 0) throw Invalid("Elastic bulk modulus must be positive");
    }
    if (!std::isfinite(vmax) || c.dt * vmax / c.h > courant_limit())
        throw Invalid("The time step exceeds the 2D staggered-grid CFL limit");
    for (const auto* p : {&x.dx, &x.dx_half, &x.dz, &x.dz_half})
        for (float d : *p)
            if (!std::isfinite(d) || d < 0) throw Invalid("PML damping must be finite and nonnegative");
    for (int r : x.rcv)
        if (r < 0 || r >= c.total_grid) throw Invalid("A receiver lies outside the simulation grid");
}

// generateRickerWavelet (src/Seis2D/Numerics2D.inc:3-14).
inline std::vector<float> ricker_wavelet(int nt, float dt, float f0, float t0) {
    std::vector<float> w(std::size_t(std::max(0, nt)));
    for (int i = 0; i < nt; ++i) {
        const float t = i * dt - t0;
        const float pi2 = float(3.141592653589793 * 3.141592653589793 * f0 * f0);
        w[std::size_t(i)] = (1.0f - 2.0f * pi2 * t * t) * std::exp(-pi2 * t * t);
    }
    return w;
}

// updatePmlDampingArrays (Numerics2D.inc:36-83) for a given maximum vp.
inline void pml_profiles(int nx, int nz, int npml, float h, float max_vp, float* dx, float* dxh, float* dz, float* dzh) {
    std::fill(dx, dx + nx, 0.0f); std::fill(dxh, dxh + nx, 0.0f);
    std::fill(dz, dz + nz, 0.0f); std::fill(dzh, dzh + nz, 0.0f);
    if (npml <= 0) return;
    if (max_vp <= 350.0f) max_vp = 2000.0f;
    const float L = npml * h;
    const float dmax = (3.0f * max_vp * 16.12f * 3.5f) / (2.0f * L);
    for (int axis = 0; axis < 2; ++axis) {
        const int n = axis ? nz : nx;
        float* w = axis ? dz : dx;
        float* hw = axis ? dzh : dxh;
        for (int i = 0; i < npml; ++i) {
            const float thick = (npml - i) / float(npml);
            float thick_h = (npml - (i + 0.5f)) / float(npml);
            if (thick_h < 0.0f) thick_h = 0.0f;
            const float v = dmax * std::pow(thick, 2.5f), vh = dmax * std::pow(thick_h, 2.5f);
            w[i] = v; w[n - 1 - i] = v; hw[i] = vh; hw[n - 1 - i] = vh;
        }
    }
}

template <class Real>
struct Solver2D {
    Context x;
    int nx = 0, nz = 0, fs = 0;
    bool up = false;
    Real inv_dt = 0, dt = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0;
    std::vector<Real> rho, mu, lambda, l2m, dp, dx, dxh, dz, dzh;
    // split fields (types 1, 3): 0 vx1 1 vx2 2 vz1 3 vz2 4 sxx1 5 sxx2 6 szz1 7 szz2 8 sxz1 9 sxz2
    // unsplit (type 2):          10 vx 11 vz 12 sxx 13 szz 14 sxz
    std::vector<Real> f[15];
    // type 2 averaged materials (Cuda_Check.cu:924-947)
    std::vector<Real> mu_x, lam_x, l2m_x, mu_z, rho_xz, rho_o;
    std::vector<seis2d_source> sources;

    explicit Solver2D(const Context& ctx) : x(ctx) {
        validate(x);
        const seis2d_ctx& c = x.c;
        nx = c.nx; nz = c.nz; fs = c.npml; up = c.up_flag != 0;
        dt = Real(c.dt);
        inv_dt = Real(1.0f / c.dt);   // host float division (Cuda_Check.cu:893)
        c1 = Real(c.c1_h); c2 = Real(c.c2_h); c3 = Real(c.c3_h); c4 = Real(c.c4_h);
        auto cvt = [](const std::vector<float>& v) { return std::vector<Real>(v.begin(), v.end()); };
        rho = cvt(x.rho); mu = cvt(x.mu); lambda = cvt(x.lambda); l2m = cvt(x.lambda2mu); dp = cvt(x.dp_flat);
        dx = cvt(x.dx); dxh = cvt(x.dx_half); dz = cvt(x.dz); dzh = cvt(x.dz_half);
        const std::size_t n = std::size_t(nx) * nz;
        for (auto& v : f) v.assign(n, Real(0));
        if (c.flag_type == 2) {
            mu_x.assign(n, 0); lam_x.assign(n, 0); l2m_x.assign(n, 0); mu_z.assign(n, 0); rho_xz.assign(n, 0);
            rho_o.assign(n, 0);
            // Averages are formed in float on the host, as the reference does.
            for (int i = 0; i < nz; ++i)
                for (int j = 0; j < nx; ++j) {
                    const int k = i * nx + j;
                    const int kx = j < nx - 1 ? k + 1 : k, kz = i < nz - 1 ? k + nx : k;
                    const int kd = (i < nz - 1 && j < nx - 1) ? k + nx + 1 : k;
                    rho_o[k] = Real(x.rho[k]);
                    mu_x[k] = Real(0.5f * (x.mu[k] + x.mu[kx]));
                    lam_x[k] = Real(0.5f * (x.lambda[k] + x.lambda[kx]));
                    l2m_x[k] = Real(0.5f * (x.lambda2mu[k] + x.lambda2mu[kx]));
                    mu_z[k] = Real(0.5f * (x.mu[k] + x.mu[kz]));
                    rho_xz[k] = Real(0.25f * (x.rho[k] + x.rho[kx] + x.rho[kz] + x.rho[kd]));
                }
        }
    }

    // S(i) - S(j) of a split pair, evaluated as ((a[i] + b[i]) - a[j]) - b[j].
    Real sd(const std::vector<Real>& a, const std::vector<Real>& b, int i, int j) const {
        return a[i] + b[i] - a[j] - b[j];
    }
    // Forward staggered derivative at k+1/2 (taps k+1..k+4 minus k..k-3), and backward at k-1/2.
    Real fwd(const std::vector<Real>& a, const std::vector<Real>& b, int k, int s) const {
        return c1 * sd(a, b, k + s, k) + c2 * sd(a, b, k + 2 * s, k - s) + c3 * sd(a, b, k + 3 * s, k - 2 * s) +
               c4 * sd(a, b, k + 4 * s, k - 3 * s);
    }
    Real bwd(const std::vector<Real>& a, const std::vector<Real>& b, int k, int s) const {
        return c1 * sd(a, b, k, k - s) + c2 * sd(a, b, k + s, k - 2 * s) + c3 * sd(a, b, k + 2 * s, k - 3 * s) +
               c4 * sd(a, b, k + 3 * s, k - 4 * s);
    }
    Real fwd1(const std::vector<Real>& a, int k, int s) const {
        return c1 * (a[k + s] - a[k]) + c2 * (a[k + 2 * s] - a[k - s]) + c3 * (a[k + 3 * s] - a[k - 2 * s]) +
               c4 * (a[k + 4 * s] - a[k - 3 * s]);
    }
    Real bwd1(const std::vector<Real>& a, int k, int s) const {
        return c1 * (a[k] - a[k - s]) + c2 * (a[k + s] - a[k - 2 * s]) + c3 * (a[k + 2 * s] - a[k - 3 * s]) +
               c4 * (a[k + 3 * s] - a[k - 4 * s]);
    }
    static Real pml(Real old, Real inv, Real d, Real src) { return (old * (inv - d) + src) / (inv + d); }

    // updateType1StressKernel (Cuda_Check.cu:225-266).
    void type1_stress() {
        auto &vx1 = f[0], &vx2 = f[1], &vz1 = f[2], &vz2 = f[3];
        #pragma omp parallel for schedule(static)
        for (int i = 4; i < nz - 4; ++i)
            for (int j = 4; j < nx - 4; ++j) {
                const int k = i * nx + j;
                const Real dvx_dx = fwd(vx1, vx2, k, 1), dvz_dz = bwd(vz1, vz2, k, nx);
                const Real l2 = Real(0.5) * (l2m[k] + l2m[k + 1]), la = Real(0.5) * (lambda[k] + lambda[k + 1]);
                f[4][k] = pml(f[4][k], inv_dt, dxh[j], l2 * dvx_dx);
                f[5][k] = pml(f[5][k], inv_dt, dz[i], la * dvz_dz);
                f[6][k] = pml(f[6][k], inv_dt, dxh[j], la * dvx_dx);
                f[7][k] = pml(f[7][k], inv_dt, dz[i], l2 * dvz_dz);
                const Real dvz_dx = bwd(vz1, vz2, k, 1), dvx_dz = fwd(vx1, vx2, k, nx);
                const Real m = Real(0.5) * (mu[k] + mu[k + nx]);
                f[8][k] = pml(f[8][k], inv_dt, dx[j], m * dvz_dx);
                f[9][k] = pml(f[9][k], inv_dt, dzh[i], m * dvx_dz);
            }
    }

    // updateType1VelocityKernel (Cuda_Check.cu:270-311).
    void type1_velocity() {
        #pragma omp parallel for schedule(static)
        for (int i = 4; i < nz - 4; ++i)
            for (int j = 4; j < nx - 4; ++j) {
                const int k = i * nx + j;
                const Real r = rho[k];
                const Real sxx_x = bwd(f[4], f[5], k, 1), sxz_z = bwd(f[8], f[9], k, nx);
                f[0][k] = pml(f[0][k], inv_dt, dx[j], sxx_x / r);
                f[1][k] = pml(f[1][k], inv_dt, dz[i], sxz_z / r);
                const Real sxz_x = fwd(f[8], f[9], k, 1), szz_z = fwd(f[6], f[7], k, nx);
                const Real ra = Real(0.25) * (rho[k] + rho[k + 1] + rho[k + nx] + rho[k + nx + 1]);
                f[2][k] = pml(f[2][k], inv_dt, dxh[j], sxz_x / ra);
                f[3][k] = pml(f[3][k], inv_dt, dzh[i], szz_z / ra);
            }
    }

    bool in_pml(int i, int j) const { return i < fs || i >= nz - fs || j < fs || j >= nx - fs; }

    // updateType2StressKernel (Cuda_Check.cu:315-360).
    void type2_stress() {
        #pragma omp parallel for schedule(static)
        for (int i = 4; i < nz - 4; ++i)
            for (int j = 4; j < nx - 4; ++j) {
                const int k = i * nx + j;
                const Real dvz_dx = bwd1(f[11], k, 1), dvx_dz = fwd1(f[10], k, nx);
                const Real dvx_dx = fwd1(f[10], k, 1), dvz_dz = bwd1(f[11], k, nx);
                if (!in_pml(i, j)) {
                    f[14][k] += dt * mu_z[k] * (dvz_dx + dvx_dz);
                    f[12][k] += dt * (l2m_x[k] * dvx_dx + lam_x[k] * dvz_dz);
                    f[13][k] += dt * (lam_x[k] * dvx_dx + l2m_x[k] * dvz_dz);
                } else {
                    f[8][k] = pml(f[8][k], inv_dt, dx[j], mu_z[k] * dvz_dx);
                    f[9][k] = pml(f[9][k], inv_dt, dzh[i], mu_z[k] * dvx_dz);
                    f[4][k] = pml(f[4][k], inv_dt, dxh[j], l2m_x[k] * dvx_dx);
                    f[6][k] = pml(f[6][k], inv_dt, dxh[j], lam_x[k] * dvx_dx);
                    f[5][k] = pml(f[5][k], inv_dt, dz[i], lam_x[k] * dvz_dz);
                    f[7][k] = pml(f[7][k], inv_dt, dz[i], l2m_x[k] * dvz_dz);
                    f[14][k] = f[8][k] + f[9][k];
                    f[12][k] = f[4][k] + f[5][k];
                    f[13][k] = f[6][k] + f[7][k];
                }
            }
    }

    // updateType2VelocityKernel (Cuda_Check.cu:364-412).
    void type2_velocity() {
        #pragma omp parallel for schedule(static)
        for (int i = 4; i < nz - 4; ++i)
            for (int j = 4; j < nx - 4; ++j) {
                const int k = i * nx + j;
                const Real sxz_x = fwd1(f[14], k, 1), szz_z = fwd1(f[13], k, nx);
                const Real sxx_x = bwd1(f[12], k, 1), sxz_z = bwd1(f[14], k, nx);
                if (!in_pml(i, j)) {
                    f[11][k] += dt * (sxz_x + szz_z) / rho_xz[k];
                    f[10][k] += dt * (sxx_x + sxz_z) / rho_o[k];
                } else {
                    f[2][k] = pml(f[2][k], inv_dt, dxh[j], sxz_x / rho_xz[k]);
                    f[3][k] = pml(f[3][k], inv_dt, dzh[i], szz_z / rho_xz[k]);
                    f[0][k] = pml(f[0][k], inv_dt, dx[j], sxx_x / rho_o[k]);
                    f[1][k] = pml(f[1][k], inv_dt, dz[i], sxz_z / rho_o[k]);
                    f[10][k] = f[0][k] + f[1][k];
                    f[11][k] = f[2][k] + f[3][k];
                }
            }
    }

    // updateType3StressKernel (Cuda_Check.cu:416-487).
    void type3_stress() {
        #pragma omp parallel for schedule(static)
        for (int i = fs; i < nz - 4; ++i)
            for (int j = 4; j < nx - 4; ++j) {
                const int k = i * nx + j;
                const bool near = i <= fs + 5 && up;
                const Real dzv = near ? Real(0) : dz[i], dzhv = near ? Real(0) : dzh[i];
                if (i == fs) {
                    f[6][k] = 0; f[7][k] = 0;
                    const Real dvx_dx = fwd(f[0], f[1], k, 1);
                    f[4][k] = pml(f[4][k], inv_dt, dxh[j], dp[j] * dvx_dx);
                    f[5][k] = 0;
                    const Real dvz_dx = c1 * sd(f[2], f[3], k, k - 1) + c2 * sd(f[2], f[3], k + 1, k - 2);
                    const Real dvx_dz = c1 * sd(f[0], f[1], k + nx, k) + c2 * sd(f[0], f[1], k + 2 * nx, k - nx);
                    const Real m = Real(0.5) * (mu[k] + mu[k + nx]);
                    f[8][k] = pml(f[8][k], inv_dt, dx[j], m * dvz_dx);
                    f[9][k] = pml(f[9][k], inv_dt, dzhv, m * dvx_dz);
                } else {
                    const Real dvx_dx = fwd(f[0], f[1], k, 1), dvz_dz = bwd(f[2], f[3], k, nx);
                    const Real l2 = Real(0.5) * (l2m[k] + l2m[k + 1]), la = Real(0.5) * (lambda[k] + lambda[k + 1]);
                    f[4][k] = pml(f[4][k], inv_dt, dxh[j], l2 * dvx_dx);
                    f[5][k] = pml(f[5][k], inv_dt, dzv, la * dvz_dz);
                    f[6][k] = pml(f[6][k], inv_dt, dxh[j], la * dvx_dx);
                    f[7][k] = pml(f[7][k], inv_dt, dzv, l2 * dvz_dz);
                    const Real dvz_dx = bwd(f[2], f[3], k, 1), dvx_dz = fwd(f[0], f[1], k, nx);
                    const Real m = Real(0.5) * (mu[k] + mu[k + nx]);
                    f[8][k] = pml(f[8][k], inv_dt, dx[j], m * dvz_dx);
                    f[9][k] = pml(f[9][k], inv_dt, dzhv, m * dvx_dz);
                }
            }
    }

    // mirrorType3StressKernel (Cuda_Check.cu:491-508).
    void type3_mirror() {
        const int n = nx * nz;
        for (int j = 4; j < nx - 4; ++j)
            for (int l = 1; l <= 4; ++l) {
                const int above = (fs - l) * nx + j, below = (fs + l) * nx + j;
                if (above >= 0 && below < n) {
                    f[6][above] = -f[6][below];
                    f[7][above] = -f[7][below];
                    const int shear = (fs + l - 1) * nx + j;
                    f[8][above] = -f[8][shear];
                    f[9][above] = -f[9][shear];
                }
            }
    }

    // updateType3VelocityKernel (Cuda_Check.cu:512-563).
    void type3_velocity() {
        #pragma omp parallel for schedule(static)
        for (int i = 4; i < nz - 4; ++i)
            for (int j = 4; j < nx - 4; ++j) {
                const int k = i * nx + j;
                if (i < fs) {
                    f[0][k] = 0; f[1][k] = 0; f[2][k] = 0; f[3][k] = 0;
                    continue;
                }
                const bool near = i <= fs + 5 && up;
                const Real dzv = near ? Real(0) : dz[i], dzhv = near ? Real(0) : dzh[i];
                const Real r = rho[k];
                const Real sxx_x = bwd(f[4], f[5], k, 1), sxz_z = bwd(f[8], f[9], k, nx);
                f[0][k] = pml(f[0][k], inv_dt, dx[j], sxx_x / r);
                f[1][k] = pml(f[1][k], inv_dt, dzv, sxz_z / r);
                const Real sxz_x = fwd(f[8], f[9], k, 1), szz_z = fwd(f[6], f[7], k, nx);
                const Real ra = Real(0.25) * (rho[k] + rho[k + 1] + rho[k + nx] + rho[k + nx + 1]);
                f[2][k] = pml(f[2][k], inv_dt, dxh[j], sxz_x / ra);
                f[3][k] = pml(f[3][k], inv_dt, dzhv, szz_z / ra);
            }
    }

    // applyPointSource (Cuda_Check.cu:567-640); sin/cos of the rotated force
    // come from the host, as in this project's launcher.
    void point_source(Real value, int idx, int type, float sin_a, float cos_a) {
        const int n = nx * nz, t = x.c.flag_type;
        if (idx < 0 || idx >= n || !std::isfinite(double(value))) return;
        const int sx = idx % nx, sz = idx / nx;
        if (sx < kRadius || sx >= nx - kRadius || sz < kRadius || sz >= nz - kRadius || (t == 3 && up && sz <= fs)) return;
        const bool split = t == 1 || t == 3;
        auto rho_avg = [&](int k) { return Real(0.25) * (rho[k] + rho[k + 1] + rho[k + nx] + rho[k + nx + 1]); };
        if (type == 0) {
            if (split) f[2][idx] += value * dt / rho_avg(idx);
            else f[11][idx] += value * dt / rho_xz[idx];
        } else if (type == 1) {
            if (split) f[0][idx] += value * dt / rho[idx];
            else f[10][idx] += value * dt / rho_o[idx];
        } else if (type == 2) {
            if (split) {
                f[4][idx] += value * Real(0.5); f[5][idx] += value * Real(0.5);
                f[6][idx] += value * Real(0.5); f[7][idx] += value * Real(0.5);
            } else {
                f[12][idx] += value; f[13][idx] += value;
            }
        } else if (type == 3) {
            if (split) { f[8][idx] += value * Real(0.5); f[9][idx] += value * Real(0.5); }
            else f[14][idx] += value;
        } else if (type == 4) {
            const Real fx = Real(sin_a) * value, fz = Real(cos_a) * value;
            if (split) {
                f[0][idx] += fx * dt / rho[idx];
                f[2][idx] += fz * dt / rho_avg(idx);
            } else {
                f[10][idx] += fx * dt / rho_o[idx];
                f[11][idx] += fz * dt / rho_xz[idx];
            }
        }
    }

    // zeroBoundaryKernel (Cuda_Check.cu:643-667).
    void zero_boundary() {
        const bool split = x.c.flag_type == 1 || x.c.flag_type == 3;
        auto zero = [&](int k) {
            if (split) { f[0][k] = 0; f[1][k] = 0; f[2][k] = 0; f[3][k] = 0; }
            else { f[10][k] = 0; f[11][k] = 0; }
        };
        for (int j = 0; j < nx; ++j) { zero(j); zero((nz - 1) * nx + j); }
        for (int i = 0; i < nz; ++i) { zero(i * nx); zero(i * nx + nx - 1); }
    }

    // runGpuSimulationStep (Cuda_Check.cu:1027-1085).
    void step(int it, int active_sources) {
        const seis2d_ctx& c = x.c;
        if (it < 0 || active_sources > kMaxSources) throw Invalid("Invalid GPU simulation state, iteration, or source count");
        const int t = c.flag_type;
        if (t == 1 || (t == 3 && !up)) type1_stress();
        else if (t == 2) type2_stress();
        else { type3_stress(); if (up) type3_mirror(); }
        if (t == 1 || (t == 3 && !up)) type1_velocity();
        else if (t == 2) type2_velocity();
        else type3_velocity();
        // sinf/cosf of float(angle * pi / 180) as the reference's source kernel (Cuda_Check.cu:622-623).
        const float ang = float(c.src_angle * 3.141592653589793 / 180.0f);
        const float sa = std::sin(ang), ca = std::cos(ang);
        if (active_sources >= 0) {
            for (int r = 0; r < active_sources && r < int(sources.size()); ++r) {
                const float v = multi_source_value(sources[std::size_t(r)]);
                if (!std::isnan(v)) point_source(Real(v * kSourceScale), sources[std::size_t(r)].idx, c.src_type, sa, ca);
            }
        } else if (it < c.nt && it < int(x.wavelet.size())) {
            const float v = x.wavelet[std::size_t(it)] * c.src_amp * kSourceScale;
            const int z = (t == 3 && up) ? std::max(c.src_z_idx, fs + 1) : c.src_z_idx;
            point_source(Real(v), z * nx + (c.src_idx % nx), c.src_type, sa, ca);
        }
        zero_boundary();
    }

    // injectMultipleSourcesKernel's Ricker (Cuda_Check.cu:822-834); NaN marks a skipped source.
    static float multi_source_value(const seis2d_source& s) {
        if (!std::isfinite(s.f_peak) || s.f_peak <= 0.0f || !std::isfinite(s.t)) return std::nanf("");
        const float tp = 1.0f / s.f_peak;
        const float xx = float(3.141592653589793 * s.f_peak * (s.t - tp));
        return s.amp * (1.0f - 2.0f * xx * xx) * std::exp(-xx * xx);
    }

    // extractReceiverSamplesKernel (Cuda_Check.cu:793-818).
    void receivers(Real* vx, Real* vz) const {
        const bool split = x.c.flag_type == 1 || x.c.flag_type == 3;
        const int n = nx * nz;
        for (std::size_t r = 0; r < x.rcv.size(); ++r) {
            const int k = x.rcv[r];
            if (k < 0 || k >= n) { vx[r] = 0; vz[r] = 0; continue; }
            vx[r] = split ? f[0][k] + f[1][k] : f[10][k];
            vz[r] = split ? f[2][k] + f[3][k] : f[11][k];
        }
    }
};

}  // namespace seis2d_oracle
