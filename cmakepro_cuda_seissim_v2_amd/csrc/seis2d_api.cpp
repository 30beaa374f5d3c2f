// seis2d_api.cpp - C ABI and host session of the 2D P-SV solver.
//
// Restates the reference's 2D host API (CUDA_Kernels/Cuda_Check.cu:841-1169)
// and host numerics (Main_header/SolverNumerics.h, src/Seis2D/Numerics2D.inc)
// on HIP, with one non-blocking stream per session.  The per-step entry
// points keep the reference's synchronous semantics; seis2d_run replays a
// captured chunk of steps and records receivers on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/seis2d.h"
#include "seis2d_device.hpp"
#include "seis2d_host.hpp"

using namespace seis2d::host;

namespace {

constexpr int kRadius = 4;                 // SolverNumerics.h:9
constexpr int kMaxSources = 200;           // SolverNumerics.h:10
constexpr float kSourceScale = 1e7f;       // SolverNumerics.h:12
constexpr int kChunk = 64;                 // steps per captured graph

void check(hipError_t e, const char* op) {
    if (e != hipSuccess) fail(SEIS2D_EHIP, std::string("HIP failure: ") + op + " | " + hipGetErrorString(e));
}

float courant_limit() {
    const double s = 1225.0 / 1024 + 245.0 / 3072 + 49.0 / 5120 + 5.0 / 7168;
    return float(1.0 / (1.4142135623730951 * s));
}

std::size_t grid_count(int nx, int nz) {
    if (nx <= 2 * kRadius || nz <= 2 * kRadius) invalid("Grid dimensions must exceed the 8-cell stencil guard region");
    const std::size_t n = std::size_t(nx) * std::size_t(nz);
    if (n > std::size_t(std::numeric_limits<int>::max())) invalid("The 2D solver index capacity was exceeded");
    return n;
}

void boundary_geometry(int nx, int nz, int pml, bool free_surface) {
    grid_count(nx, nz);
    if (pml < 0 || pml > (nx - 2 * kRadius - 1) / 2 || pml > (nz - 2 * kRadius - 1) / 2)
        invalid("PML width leaves no physical interior grid");
    if (free_surface && pml < kRadius) invalid("A free surface requires at least 4 top ghost rows");
}

// validateSimulationContext (Cuda_Check.cu:841-884).
void validate(const seis2d_ctx& c, const seis2d_arrays& a) {
    const std::size_t n = grid_count(c.nx, c.nz);
    boundary_geometry(c.nx, c.nz, c.npml, c.flag_type == 3 && c.up_flag);
    if (c.flag_type < 1 || c.flag_type > 3 || std::size_t(c.total_grid) != n || !std::isfinite(c.dt) || c.dt <= 0 ||
        !std::isfinite(c.h) || c.h <= 0 || c.nt <= 0)
        invalid("Invalid solver, time grid, or spatial grid");
    if (!a.rho || !a.mu || !a.lambda || !a.lambda2mu || !a.dx || !a.dx_half || !a.dz || !a.dz_half ||
        (c.flag_type == 3 && !a.dp_flat) || !a.wavelet || a.wavelet_len < c.nt || c.num_rcv < 0 ||
        (c.num_rcv > 0 && !a.rcv_grid_idx))
        invalid("Simulation array lengths do not match the grid or receiver count");
    float vmax = 0;
    for (std::size_t k = 0; k < n; ++k) {
        if (!std::isfinite(a.rho[k]) || a.rho[k] <= 0 || !std::isfinite(a.mu[k]) || a.mu[k] < 0 ||
            !std::isfinite(a.lambda[k]) || !std::isfinite(a.lambda2mu[k]) || a.lambda2mu[k] <= 0)
            invalid("Density/stiffness must be finite with rho > 0, mu >= 0, lambda2mu > 0");
        vmax = std::max(vmax, std::sqrt(a.lambda2mu[k] / a.rho[k]));
        if (a.lambda2mu[k] - (4.0f / 3.0f) * a.mu[k] <= 0) invalid("Elastic bulk modulus must be positive");
    }
    if (!std::isfinite(vmax) || c.dt * vmax / c.h > courant_limit())
        invalid("The time step exceeds the 2D staggered-grid CFL limit");
    const float* prof[4] = {a.dx, a.dx_half, a.dz, a.dz_half};
    const int len[4] = {c.nx, c.nx, c.nz, c.nz};
    for (int p = 0; p < 4; ++p)
        for (int i = 0; i < len[p]; ++i)
            if (!std::isfinite(prof[p][i]) || prof[p][i] < 0) invalid("PML damping must be finite and nonnegative");
    for (int r = 0; r < c.num_rcv; ++r)
        if (a.rcv_grid_idx[r] < 0 || a.rcv_grid_idx[r] >= c.total_grid)
            invalid("A receiver lies outside the simulation grid");
}

// sinf/cosf of float(angle * pi / 180), as the reference's rotated force (Cuda_Check.cu:622-623).
void angle_terms(float angle, float& s, float& c) {
    const float a = float(angle * 3.141592653589793 / 180.0f);
    s = std::sin(a);
    c = std::cos(a);
}

// injectMultipleSourcesKernel's Ricker (Cuda_Check.cu:822-834); NaN marks a skipped source.
float multi_value(const seis2d_source& s) {
    if (!std::isfinite(s.f_peak) || s.f_peak <= 0.0f || !std::isfinite(s.t)) return std::numeric_limits<float>::quiet_NaN();
    const float tp = 1.0f / s.f_peak;
    const float x = float(3.141592653589793 * s.f_peak * (s.t - tp));
    return s.amp * (1.0f - 2.0f * x * x) * std::exp(-x * x);
}

}  // namespace

struct seis2d_session {
    seis2d_ctx c{};
    hipStream_t stream = nullptr;
    std::vector<float*> owned;
    seis2d::Dev2D g{};
    int* rcv = nullptr;
    int rcv_capacity = 0;
    float* rec_vx = nullptr;
    float* rec_vz = nullptr;
    std::vector<seis2d_source> sources;
    seis2d::MultiSrc* dev_sources = nullptr;
    seis2d::MultiSrc* host_sources = nullptr;   // pinned staging
    int* counter = nullptr;

    float* alloc(std::size_t count) {
        float* p = nullptr;
        check(hipMalloc(&p, std::max<std::size_t>(1, count) * sizeof(float)), "allocate");
        owned.push_back(p);
        check(hipMemsetAsync(p, 0, std::max<std::size_t>(1, count) * sizeof(float), stream), "clear");
        return p;
    }
    float* upload(const float* src, std::size_t count) {
        float* p = alloc(count);
        check(hipMemcpyAsync(p, src, count * sizeof(float), hipMemcpyHostToDevice, stream), "upload");
        return p;
    }
    void release() noexcept {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto* p : owned) (void)hipFree(p);
        owned.clear();
        (void)hipFree(rcv); (void)hipFree(rec_vx); (void)hipFree(rec_vz); (void)hipFree(dev_sources); (void)hipFree(counter);
        if (host_sources) (void)hipHostFree(host_sources);
        rcv = nullptr; rec_vx = rec_vz = nullptr; dev_sources = nullptr; host_sources = nullptr; counter = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

namespace {

void require(const seis2d_t* t) {
    if (!t || !t->stream) fail(SEIS2D_ELIFECYCLE, "The 2D simulation is not initialized");
}

// updateGpuReceivers (Cuda_Check.cu:992-1024): prepare, then swap.
void set_receivers(seis2d_t* t, const int32_t* idx, int count) {
    if (count < 0 || (count > 0 && !idx)) invalid("Too many receivers");
    for (int r = 0; r < count; ++r)
        if (idx[r] < 0 || idx[r] >= t->g.n) invalid("Receiver index is outside the active GPU grid");
    int* ni = nullptr;
    float *nvx = nullptr, *nvz = nullptr;
    try {
        if (count > 0) {
            check(hipMalloc(&ni, std::size_t(count) * sizeof(int)), "allocate receivers");
            check(hipMemcpyAsync(ni, idx, std::size_t(count) * sizeof(int), hipMemcpyHostToDevice, t->stream),
                  "upload receivers");
            check(hipMalloc(&nvx, std::size_t(count) * sizeof(float)), "allocate receivers");
            check(hipMalloc(&nvz, std::size_t(count) * sizeof(float)), "allocate receivers");
            check(hipMemsetAsync(nvx, 0, std::size_t(count) * sizeof(float), t->stream), "clear receivers");
            check(hipMemsetAsync(nvz, 0, std::size_t(count) * sizeof(float), t->stream), "clear receivers");
        }
        check(hipStreamSynchronize(t->stream), "update receivers");
    } catch (...) {
        (void)hipFree(ni); (void)hipFree(nvx); (void)hipFree(nvz);
        throw;
    }
    (void)hipFree(t->rcv); (void)hipFree(t->rec_vx); (void)hipFree(t->rec_vz);
    t->rcv = ni; t->rec_vx = nvx; t->rec_vz = nvz;
    t->rcv_capacity = count;
}

// Single-source term of runGpuSimulationStep (Cuda_Check.cu:1070-1078).
seis2d::Src2D single_source(const seis2d_t* t, const seis2d_ctx& c, int it, const float* wavelet, int len) {
    seis2d::Src2D s{};
    const bool up = t->g.type == 3 && t->g.up;
    const int z = up ? std::max(c.src_z_idx, t->g.fs + 1) : c.src_z_idx;
    s.idx = z * t->g.nx + (c.src_idx % t->g.nx);
    s.kind = c.src_type;
    angle_terms(c.src_angle, s.sin_a, s.cos_a);
    s.active = (it < c.nt && it < len && wavelet) ? 1 : 0;
    s.value = s.active ? wavelet[it] * c.src_amp * kSourceScale : 0.0f;
    return s;
}

// One step: the stress half-step (with the type-3 mirror), the velocity half-step, then one
// workgroup for the sources, the zero frame and, when `rec` is given, the receiver samples.
void enqueue_step(seis2d_t* t, const seis2d::Dev2D& g, const seis2d::Src2D& s, int active_sources, const seis2d_ctx& c,
                  const seis2d::Rec2D* rec = nullptr) {
    check(seis2d::launch_stress2d(g, t->stream), "kernel launch");
    seis2d::Tail2D tail{};
    tail.single = s;
    if (active_sources >= 0) {
        tail.single.active = 0;
        if (active_sources > 0) {
            tail.list = t->dev_sources;
            tail.count = active_sources;
            tail.kind = c.src_type;
            angle_terms(c.src_angle, tail.sin_a, tail.cos_a);
        }
    }
    if (rec) tail.rec = *rec;
    check(seis2d::launch_velocity2d(g, tail, t->stream), "kernel launch");
}

void record_step(seis2d_t* t, int num_rcv) {
    seis2d::Rec2D r{t->rcv, num_rcv, t->rec_vx, t->rec_vz, 0, 0};
    check(seis2d::launch_receivers2d(t->g, r, t->stream), "kernel launch");
}

float* field_ptr(seis2d_t* t, int field) {
    if (field < 0 || field >= 15 || !t->g.f[field]) invalid("Unknown or unallocated 2D field");
    return t->g.f[field];
}

}  // namespace

extern "C" {

float seis2d_courant_limit(void) { return courant_limit(); }

int seis2d_grid_count(int nx, int nz, int64_t* count) {
    return guarded([&] {
        const std::size_t n = grid_count(nx, nz);
        if (count) *count = int64_t(n);
    });
}

int seis2d_boundary_geometry(int nx, int nz, int npml, int free_surface) {
    return guarded([&] { boundary_geometry(nx, nz, npml, free_surface != 0); });
}

int seis2d_sampling_stride(double dt, double rate_hz, int* stride) {
    return guarded([&] {
        if (!std::isfinite(dt) || dt <= 0.0 || !std::isfinite(rate_hz) || rate_hz <= 0.0)
            invalid("Sampling rate and simulation time step must be finite and positive");
        const double s = std::round(1.0 / (rate_hz * dt));
        if (!std::isfinite(s) || s > double(std::numeric_limits<int>::max())) invalid("Receiver sampling stride is too large");
        if (stride) *stride = s < 1.0 ? 1 : int(s);
    });
}

int seis2d_ricker_wavelet(int nt, float dt, float f0, float t0, float* out) {
    return guarded([&] {
        if (nt <= 0 || !std::isfinite(dt) || dt <= 0.0f || !std::isfinite(f0) || f0 <= 0.0f || !std::isfinite(t0) || !out)
            invalid("Invalid Ricker wavelet sampling or frequency");
        for (int i = 0; i < nt; ++i) {
            const float tt = i * dt - t0;
            const float pi2 = float(3.141592653589793 * 3.141592653589793 * f0 * f0);
            out[i] = (1.0f - 2.0f * pi2 * tt * tt) * std::exp(-pi2 * tt * tt);
        }
    });
}

int seis2d_pml_profiles(const seis2d_ctx* ctx, const float* l2m, const float* rho, float* dx, float* dxh, float* dz,
                        float* dzh) {
    return guarded([&] {
        if (!ctx || !l2m || !rho || !dx || !dxh || !dz || !dzh) invalid("Invalid PML profile request");
        const seis2d_ctx& c = *ctx;
        boundary_geometry(c.nx, c.nz, c.npml, c.flag_type == 3 && c.up_flag);
        const std::size_t n = std::size_t(c.nx) * c.nz;
        float vmax = 0.0f;
        for (std::size_t k = 0; k < n; ++k) {
            const float v = std::sqrt(l2m[k] / rho[k]);
            if (v > vmax) vmax = v;
        }
        if (vmax <= 350.0f) vmax = 2000.0f;
        std::fill(dx, dx + c.nx, 0.0f); std::fill(dxh, dxh + c.nx, 0.0f);
        std::fill(dz, dz + c.nz, 0.0f); std::fill(dzh, dzh + c.nz, 0.0f);
        if (c.npml <= 0) return;
        const float L = c.npml * c.h;
        const float dmax = (3.0f * vmax * 16.12f * 3.5f) / (2.0f * L);
        for (int axis = 0; axis < 2; ++axis) {
            const int len = axis ? c.nz : c.nx;
            float* w = axis ? dz : dx;
            float* hw = axis ? dzh : dxh;
            for (int i = 0; i < c.npml; ++i) {
                const float thick = (c.npml - i) / float(c.npml);
                float thick_h = (c.npml - (i + 0.5f)) / float(c.npml);
                if (thick_h < 0.0f) thick_h = 0.0f;
                const float v = dmax * std::pow(thick, 2.5f), vh = dmax * std::pow(thick_h, 2.5f);
                w[i] = v; w[len - 1 - i] = v; hw[i] = vh; hw[len - 1 - i] = vh;
            }
        }
    });
}

int seis2d_validate(const seis2d_ctx* ctx, const seis2d_arrays* a) {
    return guarded([&] {
        if (!ctx || !a) invalid("Invalid solver, time grid, or spatial grid");
        validate(*ctx, *a);
    });
}

int seis2d_create(const seis2d_ctx* ctx, const seis2d_arrays* a, seis2d_t** out) {
    return guarded([&] {
        if (!ctx || !a || !out) invalid("Invalid solver, time grid, or spatial grid");
        validate(*ctx, *a);
        auto* t = new seis2d_session();
        try {
            const seis2d_ctx& c = *ctx;
            t->c = c;
            check(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking), "create stream");
            const std::size_t n = std::size_t(c.nx) * c.nz;
            seis2d::Dev2D& g = t->g;
            g.nx = c.nx; g.nz = c.nz; g.n = int(n);
            g.fs = c.npml; g.up = c.up_flag ? 1 : 0; g.type = c.flag_type;
            g.dt = c.dt; g.inv_dt = 1.0f / c.dt;
            g.c1 = c.c1_h; g.c2 = c.c2_h; g.c3 = c.c3_h; g.c4 = c.c4_h;
            for (int k = 0; k < 10; ++k) g.f[k] = t->alloc(n);
            if (c.flag_type == 2)
                for (int k = 10; k < 15; ++k) g.f[k] = t->alloc(n);
            g.rho = t->upload(a->rho, n); g.mu = t->upload(a->mu, n);
            g.lambda = t->upload(a->lambda, n); g.l2m = t->upload(a->lambda2mu, n);
            g.dx = t->upload(a->dx, std::size_t(c.nx)); g.dxh = t->upload(a->dx_half, std::size_t(c.nx));
            g.dz = t->upload(a->dz, std::size_t(c.nz)); g.dzh = t->upload(a->dz_half, std::size_t(c.nz));
            if (c.flag_type == 3) g.dp = t->upload(a->dp_flat, std::size_t(c.nx));
            std::vector<std::vector<float>> flats;
            if (c.flag_type == 2) {
                // Face-averaged materials, formed in float on the host (Cuda_Check.cu:924-947).
                flats.assign(6, std::vector<float>(n, 0.0f));
                for (int i = 0; i < c.nz; ++i)
                    for (int j = 0; j < c.nx; ++j) {
                        const int k = i * c.nx + j;
                        const int kx = j < c.nx - 1 ? k + 1 : k, kz = i < c.nz - 1 ? k + c.nx : k;
                        const int kd = (i < c.nz - 1 && j < c.nx - 1) ? k + c.nx + 1 : k;
                        flats[0][k] = 0.5f * (a->mu[k] + a->mu[kx]);
                        flats[1][k] = 0.5f * (a->lambda[k] + a->lambda[kx]);
                        flats[2][k] = 0.5f * (a->lambda2mu[k] + a->lambda2mu[kx]);
                        flats[3][k] = 0.5f * (a->mu[k] + a->mu[kz]);
                        flats[4][k] = 0.25f * (a->rho[k] + a->rho[kx] + a->rho[kz] + a->rho[kd]);
                        flats[5][k] = a->rho[k];
                    }
                g.mu_x = t->upload(flats[0].data(), n); g.lam_x = t->upload(flats[1].data(), n);
                g.l2m_x = t->upload(flats[2].data(), n); g.mu_z = t->upload(flats[3].data(), n);
                g.rho_xz = t->upload(flats[4].data(), n); g.rho_o = t->upload(flats[5].data(), n);
            }
            check(hipMalloc(&t->dev_sources, kMaxSources * sizeof(seis2d::MultiSrc)), "allocate sources");
            check(hipHostMalloc(&t->host_sources, kMaxSources * sizeof(seis2d::MultiSrc), hipHostMallocDefault),
                  "allocate sources");
            check(hipMalloc(&t->counter, sizeof(int)), "allocate step counter");
            check(hipStreamSynchronize(t->stream), "initialize 2D simulation");
            set_receivers(t, a->rcv_grid_idx, c.num_rcv);
        } catch (...) {
            t->release();
            delete t;
            throw;
        }
        *out = t;
    });
}

void seis2d_destroy(seis2d_t* t) {
    if (!t) return;
    t->release();
    delete t;
}

int seis2d_update_receivers(seis2d_t* t, const int32_t* idx, int count) {
    return guarded([&] {
        require(t);
        set_receivers(t, idx, count);
    });
}

int seis2d_set_sources(seis2d_t* t, const seis2d_source* sources, int count) {
    return guarded([&] {
        require(t);
        if (count < 0 || count > kMaxSources || (count && !sources)) invalid("Active source capacity exceeded");
        t->sources.assign(sources, sources + count);
    });
}

int seis2d_step(seis2d_t* t, int it, const seis2d_ctx* ctx, const float* wavelet, int wavelet_len, int active_sources) {
    return guarded([&] {
        require(t);
        // runGpuSimulationStep's checks (Cuda_Check.cu:1028-1030); more sources than were uploaded is rejected too.
        if (!ctx || it < 0 || active_sources > kMaxSources || active_sources > int(t->sources.size()))
            invalid("Invalid GPU simulation state, iteration, or source count");
        const seis2d_ctx& c = *ctx;
        if (active_sources > 0) {
            // Evaluate the active sources' Ricker values on the host and stage them.
            check(hipStreamSynchronize(t->stream), "stage sources");
            int m = 0;
            for (int r = 0; r < active_sources; ++r) {
                const float v = multi_value(t->sources[std::size_t(r)]);
                if (std::isnan(v)) continue;
                t->host_sources[m++] = {t->sources[std::size_t(r)].idx, v * kSourceScale};
            }
            check(hipMemcpyAsync(t->dev_sources, t->host_sources, std::size_t(m) * sizeof(seis2d::MultiSrc),
                                 hipMemcpyHostToDevice, t->stream),
                  "upload sources");
            active_sources = m;
        }
        seis2d::Dev2D g = t->g;
        g.counter = nullptr;
        const seis2d::Src2D s = single_source(t, c, it, wavelet, wavelet_len);
        enqueue_step(t, g, s, active_sources, c);
    });
}

int seis2d_copy_wavefield(seis2d_t* t, float* vz_out) {
    return guarded([&] {
        require(t);
        if (!vz_out) invalid("Invalid wavefield output buffer");
        const std::size_t n = std::size_t(t->g.n);
        if (t->g.type == 1 || t->g.type == 3) {
            std::vector<float> a(n), b(n);
            check(hipMemcpyAsync(a.data(), t->g.f[2], n * sizeof(float), hipMemcpyDeviceToHost, t->stream), "copy vz");
            check(hipMemcpyAsync(b.data(), t->g.f[3], n * sizeof(float), hipMemcpyDeviceToHost, t->stream), "copy vz");
            check(hipStreamSynchronize(t->stream), "copy vz");
            for (std::size_t k = 0; k < n; ++k) vz_out[k] = a[k] + b[k];
        } else {
            check(hipMemcpyAsync(vz_out, t->g.f[11], n * sizeof(float), hipMemcpyDeviceToHost, t->stream), "copy vz");
            check(hipStreamSynchronize(t->stream), "copy vz");
        }
    });
}

int seis2d_record_receiver_step(seis2d_t* t, int num_rcv, float* vx, float* vz) {
    return guarded([&] {
        require(t);
        if (num_rcv <= 0) return;
        if (num_rcv > t->rcv_capacity || !vx || !vz) invalid("Invalid receiver count or host output buffer");
        record_step(t, num_rcv);
        check(hipMemcpyAsync(vx, t->rec_vx, std::size_t(num_rcv) * sizeof(float), hipMemcpyDeviceToHost, t->stream),
              "copy receivers");
        check(hipMemcpyAsync(vz, t->rec_vz, std::size_t(num_rcv) * sizeof(float), hipMemcpyDeviceToHost, t->stream),
              "copy receivers");
        check(hipStreamSynchronize(t->stream), "copy receivers");
    });
}

int seis2d_copy_receiver_records(seis2d_t* t, int num_rcv, int nt, int it, float* rec_vx, float* rec_vz) {
    return guarded([&] {
        require(t);
        if (num_rcv <= 0) return;
        if (num_rcv > t->rcv_capacity || nt <= 0 || it < 0 || it >= nt || !rec_vx || !rec_vz)
            invalid("Invalid receiver trace count, time index, or host output buffer");
        std::vector<float> vx(static_cast<std::size_t>(num_rcv)), vz(static_cast<std::size_t>(num_rcv));
        record_step(t, num_rcv);
        check(hipMemcpyAsync(vx.data(), t->rec_vx, vx.size() * sizeof(float), hipMemcpyDeviceToHost, t->stream),
              "copy receivers");
        check(hipMemcpyAsync(vz.data(), t->rec_vz, vz.size() * sizeof(float), hipMemcpyDeviceToHost, t->stream),
              "copy receivers");
        check(hipStreamSynchronize(t->stream), "copy receivers");
        for (int r = 0; r < num_rcv; ++r) {
            rec_vx[std::size_t(r) * nt + it] = vx[std::size_t(r)];
            rec_vz[std::size_t(r) * nt + it] = vz[std::size_t(r)];
        }
    });
}

int seis2d_synchronize(seis2d_t* t) {
    return guarded([&] {
        require(t);
        check(hipStreamSynchronize(t->stream), "synchronize");
    });
}

int seis2d_run(seis2d_t* t, const seis2d_ctx* ctx, const float* wavelet, int wavelet_len, int first_it, int n_steps,
               float* out_vx, float* out_vz) {
    return guarded([&] {
        require(t);
        if (!ctx || first_it < 0 || n_steps < 1 || n_steps > 100000000) invalid("Invalid 2D run request");
        const seis2d_ctx& c = *ctx;
        const int nrec = t->rcv_capacity;
        // Per-step source values: wavelet[it] * amp * 1e7, NaN where runGpuSimulationStep injects nothing.
        std::vector<float> table(std::size_t(first_it) + std::size_t(n_steps), std::numeric_limits<float>::quiet_NaN());
        for (int it = first_it; it < first_it + n_steps; ++it)
            if (wavelet && it < c.nt && it < wavelet_len) table[std::size_t(it)] = wavelet[it] * c.src_amp * kSourceScale;
        float* dev_table = nullptr;
        float *tvx = nullptr, *tvz = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        try {
            check(hipMalloc(&dev_table, table.size() * sizeof(float)), "allocate source table");
            check(hipMemcpyAsync(dev_table, table.data(), table.size() * sizeof(float), hipMemcpyHostToDevice, t->stream),
                  "upload source table");
            const std::size_t trace = std::size_t(std::max(1, nrec)) * std::size_t(n_steps);
            check(hipMalloc(&tvx, trace * sizeof(float)), "allocate traces");
            check(hipMalloc(&tvz, trace * sizeof(float)), "allocate traces");
            const int start = first_it - 1;
            check(hipMemcpyAsync(t->counter, &start, sizeof(int), hipMemcpyHostToDevice, t->stream), "set step counter");
            check(hipStreamSynchronize(t->stream), "prepare run");

            seis2d::Dev2D g = t->g;
            g.counter = t->counter;
            seis2d::Src2D s = single_source(t, c, first_it, wavelet, wavelet_len);
            s.active = 1;
            s.table = dev_table;
            const seis2d::Rec2D rec{t->rcv, nrec, tvx, tvz, n_steps, first_it};
            auto one_step = [&] { enqueue_step(t, g, s, -1, c, &rec); };
            const int chunk = std::min(kChunk, n_steps);
            check(hipStreamBeginCapture(t->stream, hipStreamCaptureModeThreadLocal), "begin capture");
            try {
                for (int k = 0; k < chunk; ++k) one_step();
            } catch (...) {
                hipGraph_t partial = nullptr;
                (void)hipStreamEndCapture(t->stream, &partial);
                if (partial) (void)hipGraphDestroy(partial);
                throw;
            }
            check(hipStreamEndCapture(t->stream, &graph), "end capture");
            check(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0), "instantiate graph");
            const int replays = n_steps / chunk;
            for (int r = 0; r < replays; ++r) check(hipGraphLaunch(exec, t->stream), "launch graph");
            for (int k = replays * chunk; k < n_steps; ++k) one_step();
            if (out_vx && out_vz && nrec > 0) {
                check(hipMemcpyAsync(out_vx, tvx, std::size_t(nrec) * n_steps * sizeof(float), hipMemcpyDeviceToHost,
                                     t->stream),
                      "copy traces");
                check(hipMemcpyAsync(out_vz, tvz, std::size_t(nrec) * n_steps * sizeof(float), hipMemcpyDeviceToHost,
                                     t->stream),
                      "copy traces");
            }
            check(hipStreamSynchronize(t->stream), "2D run");
        } catch (...) {
            if (t->stream) (void)hipStreamSynchronize(t->stream);
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipFree(dev_table); (void)hipFree(tvx); (void)hipFree(tvz);
            throw;
        }
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        (void)hipFree(dev_table); (void)hipFree(tvx); (void)hipFree(tvz);
    });
}

int seis2d_download_field(seis2d_t* t, int field, float* out) {
    return guarded([&] {
        require(t);
        if (!out) invalid("Invalid field output buffer");
        float* p = field_ptr(t, field);
        check(hipMemcpyAsync(out, p, std::size_t(t->g.n) * sizeof(float), hipMemcpyDeviceToHost, t->stream), "copy field");
        check(hipStreamSynchronize(t->stream), "copy field");
    });
}

int seis2d_upload_field(seis2d_t* t, int field, const float* in) {
    return guarded([&] {
        require(t);
        if (!in) invalid("Invalid field input buffer");
        float* p = field_ptr(t, field);
        check(hipMemcpyAsync(p, in, std::size_t(t->g.n) * sizeof(float), hipMemcpyHostToDevice, t->stream), "upload field");
        check(hipStreamSynchronize(t->stream), "upload field");
    });
}

const char* seis2d_last_error(void) { return g_error.c_str(); }

}  // extern "C"
