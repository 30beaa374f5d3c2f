// oracle2d_capi.cpp - extern "C" entry points of the 2D CPU oracle for the
// Python tests and the cpu_baseline leg of bench.py (loaded with ctypes).
//
// TEST INFRASTRUCTURE ONLY: the checker, never the product path.
#include <chrono>
#include <memory>
#include <string>

#include "seis2d_oracle.hpp"

namespace {
thread_local std::string g_error2;

struct Handle2D {
    int precision = 32;
    std::unique_ptr<seis2d_oracle::Solver2D<float>> f32;
    std::unique_ptr<seis2d_oracle::Solver2D<double>> f64;
};

template <class F>
int guard2(F&& fn) {
    try { fn(); return 0; }
    catch (const std::invalid_argument& e) { g_error2 = e.what(); return 1; }
    catch (const std::exception& e) { g_error2 = e.what(); return 2; }
}

template <class Fn>
void with2(Handle2D* h, Fn&& fn) {
    if (h->precision == 32) fn(*h->f32);
    else fn(*h->f64);
}
}  // namespace

extern "C" {

const char* oracle2d_last_error() { return g_error2.c_str(); }

int oracle2d_validate(const seis2d_ctx* c, const seis2d_arrays* a) {
    return guard2([&] { seis2d_oracle::validate(seis2d_oracle::context_from(*c, *a)); });
}

int oracle2d_ricker(int nt, float dt, float f0, float t0, float* out) {
    return guard2([&] {
        const auto w = seis2d_oracle::ricker_wavelet(nt, dt, f0, t0);
        std::copy(w.begin(), w.end(), out);
    });
}

int oracle2d_pml(int nx, int nz, int npml, float h, float max_vp, float* dx, float* dxh, float* dz, float* dzh) {
    return guard2([&] { seis2d_oracle::pml_profiles(nx, nz, npml, h, max_vp, dx, dxh, dz, dzh); });
}

int oracle2d_create(const seis2d_ctx* c, const seis2d_arrays* a, int precision, Handle2D** out) {
    return guard2([&] {
        auto h = std::make_unique<Handle2D>();
        const auto ctx = seis2d_oracle::context_from(*c, *a);
        h->precision = precision == 64 ? 64 : 32;
        if (h->precision == 32) h->f32 = std::make_unique<seis2d_oracle::Solver2D<float>>(ctx);
        else h->f64 = std::make_unique<seis2d_oracle::Solver2D<double>>(ctx);
        *out = h.release();
    });
}

void oracle2d_destroy(Handle2D* h) { delete h; }

int oracle2d_set_sources(Handle2D* h, const seis2d_source* s, int count) {
    return guard2([&] { with2(h, [&](auto& o) { o.sources.assign(s, s + count); }); });
}

// One runGpuSimulationStep with the source scalars of `ctx` (the reference reads them each step).
int oracle2d_step(Handle2D* h, int it, const seis2d_ctx* ctx, int active_sources) {
    return guard2([&] {
        with2(h, [&](auto& o) {
            if (ctx) {
                o.x.c.src_type = ctx->src_type; o.x.c.src_angle = ctx->src_angle; o.x.c.src_amp = ctx->src_amp;
                o.x.c.src_idx = ctx->src_idx; o.x.c.src_z_idx = ctx->src_z_idx; o.x.c.nt = ctx->nt;
            }
            o.step(it, active_sources);
        });
    });
}

// n steps from `first`, recording every receiver after every step into [r][n] traces.
int oracle2d_run(Handle2D* h, int first, int n, double* rec_vx, double* rec_vz) {
    return guard2([&] {
        with2(h, [&](auto& o) {
            using Real = std::decay_t<decltype(o.dt)>;
            const std::size_t nr = o.x.rcv.size();
            std::vector<Real> vx(nr), vz(nr);
            for (int k = 0; k < n; ++k) {
                o.step(first + k, -1);
                if (rec_vx && rec_vz) {
                    o.receivers(vx.data(), vz.data());
                    for (std::size_t r = 0; r < nr; ++r) {
                        rec_vx[r * std::size_t(n) + std::size_t(k)] = double(vx[r]);
                        rec_vz[r * std::size_t(n) + std::size_t(k)] = double(vz[r]);
                    }
                }
            }
        });
    });
}

double oracle2d_time_steps(Handle2D* h, int first, int n) {
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = guard2([&] { with2(h, [&](auto& o) { for (int k = 0; k < n; ++k) o.step(first + k, -1); }); });
    const auto t1 = std::chrono::steady_clock::now();
    return rc ? -1.0 : std::chrono::duration<double>(t1 - t0).count();
}

int oracle2d_get_field(Handle2D* h, int field, double* out) {
    return guard2([&] {
        with2(h, [&](auto& o) {
            if (field < 0 || field >= 15) throw seis2d_oracle::Invalid("unknown field");
            for (std::size_t k = 0; k < o.f[field].size(); ++k) out[k] = double(o.f[field][k]);
        });
    });
}

int oracle2d_set_field(Handle2D* h, int field, const double* in) {
    return guard2([&] {
        with2(h, [&](auto& o) {
            using Real = std::decay_t<decltype(o.dt)>;
            if (field < 0 || field >= 15) throw seis2d_oracle::Invalid("unknown field");
            for (std::size_t k = 0; k < o.f[field].size(); ++k) o.f[field][k] = Real(in[k]);
        });
    });
}

int oracle2d_receivers(Handle2D* h, double* vx, double* vz) {
    return guard2([&] {
        with2(h, [&](auto& o) {
            using Real = std::decay_t<decltype(o.dt)>;
            std::vector<Real> a(o.x.rcv.size()), b(o.x.rcv.size());
            o.receivers(a.data(), b.data());
            for (std::size_t r = 0; r < a.size(); ++r) { vx[r] = double(a[r]); vz[r] = double(b[r]); }
        });
    });
}

}  // extern "C"
