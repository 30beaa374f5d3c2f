// oracle_capi.cpp - extern "C" entry points of the CPU oracle for the Python
// tests and the cpu_baseline leg of bench.py (loaded with ctypes).
//
// TEST INFRASTRUCTURE ONLY: the checker, never the product path.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <exception>
#include <memory>
#include <new>
#include <string>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "seis3d_oracle.hpp"

namespace {
thread_local std::string g_error;

struct Handle {
    int precision = 32;
    std::unique_ptr<seis3d_oracle::Solver<float>> f32;
    std::unique_ptr<seis3d_oracle::Solver<double>> f64;
};

template <class F>
int guard(F&& fn) {
    try { fn(); return 0; }
    catch (const std::invalid_argument& e) { g_error = e.what(); return 1; }
    catch (const std::exception& e) { g_error = e.what(); return 2; }
}

template <class Fn>
auto with_solver(Handle* h, Fn&& fn) {
    if (h->precision == 32) return fn(*h->f32);
    return fn(*h->f64);
}
}  // namespace

extern "C" {

const char* oracle3d_last_error() { return g_error.c_str(); }

int oracle3d_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

int oracle3d_validate(seis3d_ctx* ctx) { return guard([&] { seis3d_oracle::validate(*ctx); }); }

int oracle3d_coefficients(const seis3d_ctx* ctx, int axis, float* out) {
    return guard([&] {
        const auto v = seis3d_oracle::make_coefficients<float>(*ctx, axis);
        std::copy(v.begin(), v.end(), out);
    });
}

uint64_t oracle3d_slab_count(const seis3d_ctx* ctx, int axis) { return seis3d_oracle::slab_count(*ctx, axis); }

double oracle3d_ricker(float frequency, double time) { return seis3d_oracle::ricker_double(frequency, time); }

int oracle3d_fit_survey(const seis3d_ctx* ctx, seis3d_survey* s) {
    return guard([&] { seis3d_oracle::fit_survey(*ctx, *s); });
}

int oracle3d_survey_receivers(const seis3d_ctx* ctx, const seis3d_survey* s, seis3d_receiver* out, int* count) {
    return guard([&] {
        const auto r = seis3d_oracle::survey_receivers(*ctx, *s);
        if (int(r.size()) > *count) throw std::invalid_argument("receiver buffer too small");
        std::copy(r.begin(), r.end(), out);
        *count = int(r.size());
    });
}

int oracle3d_create(const seis3d_ctx* ctx, int precision, Handle** out) {
    return guard([&] {
        auto h = std::make_unique<Handle>();
        h->precision = precision == 64 ? 64 : 32;
        if (h->precision == 32) h->f32 = std::make_unique<seis3d_oracle::Solver<float>>(*ctx);
        else h->f64 = std::make_unique<seis3d_oracle::Solver<double>>(*ctx);
        *out = h.release();
    });
}

void oracle3d_destroy(Handle* h) { delete h; }

int oracle3d_reset(Handle* h) { return guard([&] { with_solver(h, [](auto& s) { s.reset(); }); }); }

int oracle3d_set_profile(Handle* h, const double* profile) {
    return guard([&] { with_solver(h, [&](auto& s) { s.set_profile(profile); }); });
}

int oracle3d_profile_coefficients(const seis3d_ctx* ctx, const double* profile, int axis, float* out) {
    return guard([&] {
        const auto v = seis3d_oracle::make_profile_coefficients<float>(*ctx, axis, profile);
        std::copy(v.begin(), v.end(), out);
    });
}

int oracle3d_refresh_cpml(Handle* h, const seis3d_ctx* ctx) {
    return guard([&] { with_solver(h, [&](auto& s) { s.refresh_cpml(*ctx); }); });
}

int oracle3d_upload_plane(Handle* h, int z, const float* values) {
    return guard([&] { with_solver(h, [&](auto& s) { s.upload_plane(z, values); }); });
}

int oracle3d_finish_upload(Handle* h) { return guard([&] { with_solver(h, [](auto& s) { s.finish_upload(); }); }); }

int oracle3d_set_receivers(Handle* h, const seis3d_receiver* r, int count) {
    return guard([&] { with_solver(h, [&](auto& s) { s.set_receivers(r, count); }); });
}

int oracle3d_step(Handle* h, const seis3d_source* src, uint64_t step) {
    return guard([&] { with_solver(h, [&](auto& s) { s.step(*src, step); }); });
}

// n_steps steps from first_step; after each step s with (s+1) % stride == 0
// the receivers are sampled into gather[sample][r][3] (up to max_samples).
int oracle3d_run(Handle* h, const seis3d_source* src, uint64_t first_step, int64_t n_steps, int stride,
                 double* gather, int max_samples, int* samples) {
    return guard([&] {
        with_solver(h, [&](auto& s) {
            using Real = std::decay_t<decltype(s.f[0][0])>;
            std::vector<Real> buf(s.receivers.size() * 3);
            int count = 0;
            for (int64_t k = 0; k < n_steps; ++k) {
                const uint64_t step = first_step + uint64_t(k);
                s.step(*src, step);
                if (stride > 0 && gather && (step + 1) % uint64_t(stride) == 0 && count < max_samples) {
                    s.sample(buf.data());
                    for (std::size_t j = 0; j < buf.size(); ++j) gather[std::size_t(count) * buf.size() + j] = double(buf[j]);
                    ++count;
                }
            }
            if (samples) *samples = count;
        });
    });
}

// Wall-clock seconds of n_steps steps (the CPU baseline timing).
double oracle3d_time_steps(Handle* h, const seis3d_source* src, uint64_t first_step, int64_t n_steps) {
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = oracle3d_run(h, src, first_step, n_steps, 0, nullptr, 0, nullptr);
    const auto t1 = std::chrono::steady_clock::now();
    return rc ? -1.0 : std::chrono::duration<double>(t1 - t0).count();
}

int oracle3d_sample(Handle* h, double* out) {
    return guard([&] {
        with_solver(h, [&](auto& s) {
            using Real = std::decay_t<decltype(s.f[0][0])>;
            std::vector<Real> buf(s.receivers.size() * 3);
            s.sample(buf.data());
            for (std::size_t j = 0; j < buf.size(); ++j) out[j] = double(buf[j]);
        });
    });
}

// field 0..8 wavefields, 9 1/rho, 10 lambda, 11 mu; padded reference layout.
int oracle3d_get_field(Handle* h, int field, double* out) {
    return guard([&] {
        with_solver(h, [&](auto& s) {
            const auto& v = field < 9 ? s.f[field] : field == 9 ? s.inv_rho : field == 10 ? s.lambda : s.mu;
            for (std::size_t i = 0; i < v.size(); ++i) out[i] = double(v[i]);
        });
    });
}

int oracle3d_set_field(Handle* h, int field, const double* in) {
    return guard([&] {
        with_solver(h, [&](auto& s) {
            using Real = std::decay_t<decltype(s.f[0][0])>;
            auto& v = field < 9 ? s.f[field] : field == 9 ? s.inv_rho : field == 10 ? s.lambda : s.mu;
            for (std::size_t i = 0; i < v.size(); ++i) v[i] = Real(in[i]);
        });
    });
}

int oracle3d_sample_scalar(Handle* h, int x, int y, int z, int component, double* out) {
    return guard([&] { with_solver(h, [&](auto& s) { *out = double(s.sample_scalar(x, y, z, component)); }); });
}

// Misfit and gradient (vp, vs, rho on the physical grid [z][y][x]) of
// n_steps steps sampled every `stride`; observed is [sample][receiver][3].
int oracle3d_fwi_gradient(Handle* h, const seis3d_source* src, int64_t n_steps, int stride, const double* observed,
                          double* misfit, double* g_vp, double* g_vs, double* g_rho) {
    return guard([&] {
        with_solver(h, [&](auto& s) {
            using Real = std::decay_t<decltype(s.f[0][0])>;
            const std::size_t nobs = std::size_t(n_steps / stride) * s.receivers.size() * 3;
            const std::size_t nphys = std::size_t(s.c.nx) * s.c.ny * s.c.nz;
            std::vector<Real> obs(observed, observed + nobs), a(nphys), b(nphys), c(nphys);
            *misfit = double(s.fwi_gradient(*src, n_steps, stride, obs.data(), a.data(), b.data(), c.data()));
            for (std::size_t k = 0; k < nphys; ++k) { g_vp[k] = double(a[k]); g_vs[k] = double(b[k]); g_rho[k] = double(c[k]); }
        });
    });
}

// CPML memories of one axis: every slot, in the solver's slab layout.
uint64_t oracle3d_cpml_size(Handle* h, int axis) {
    return with_solver(h, [&](auto& s) { return uint64_t(s.cpml[axis].size()); });
}

int oracle3d_get_cpml(Handle* h, int axis, double* out) {
    return guard([&] {
        with_solver(h, [&](auto& s) {
            const auto& v = s.cpml[axis];
            for (std::size_t i = 0; i < v.size(); ++i) out[i] = double(v[i]);
        });
    });
}

int oracle3d_set_cpml(Handle* h, int axis, const double* in) {
    return guard([&] {
        with_solver(h, [&](auto& s) {
            using Real = std::decay_t<decltype(s.f[0][0])>;
            auto& v = s.cpml[axis];
            for (std::size_t i = 0; i < v.size(); ++i) v[i] = Real(in[i]);
        });
    });
}

// Adjoint of one step n applied in place to an adjoint state laid out as the nine
// wavefields over the padded grid followed by the three CPML memory arrays.  The
// solver's own state stands in for the stored forward states, which only enter the
// model-gradient sums (discarded here).
int oracle3d_adjoint_step(Handle* h, const seis3d_source* src, uint64_t step, double* state) {
    return guard([&] {
        with_solver(h, [&](auto& s) {
            auto A = s.make_adjoint();
            double* p = state;
            for (auto& v : A.f) for (auto& x : v) x = *p++;
            for (auto& v : A.cpml) for (auto& x : v) x = *p++;
            const auto snap = s.save();
            s.adjoint_step(snap, snap, A, *src, step);
            p = state;
            for (const auto& v : A.f) for (const auto x : v) *p++ = double(x);
            for (const auto& v : A.cpml) for (const auto x : v) *p++ = double(x);
        });
    });
}

// Host STREAM triad a[i] = b[i] + s * c[i] over n doubles on `threads` OpenMP threads,
// first-touched by the same threads; the best of `reps` in GB/s, counting 24 bytes per
// element as STREAM does.  Gives the CPU baseline a bandwidth ceiling to compare with.
double oracle3d_stream_triad(int64_t n, int threads, int reps) {
    if (n <= 0 || reps <= 0) return 0.0;
    double* a = new (std::nothrow) double[std::size_t(n)];
    double* b = new (std::nothrow) double[std::size_t(n)];
    double* c = new (std::nothrow) double[std::size_t(n)];
    double best = 0.0;
    if (a && b && c) {
#ifdef _OPENMP
        const int saved = omp_get_max_threads();
        omp_set_num_threads(std::max(1, threads));
#endif
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) { a[i] = 0.0; b[i] = 1.0; c[i] = 2.0; }
        for (int r = 0; r < reps; ++r) {
            const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) a[i] = b[i] + 3.0 * c[i];
            const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (s > 0) best = std::max(best, 24.0 * double(n) / s / 1e9);
        }
#ifdef _OPENMP
        omp_set_num_threads(saved);
#endif
    }
    delete[] a;
    delete[] b;
    delete[] c;
    return best;
}

int oracle3d_nonfinite(Handle* h) {
    return with_solver(h, [](auto& s) { return s.nonfinite ? 1 : 0; });
}

}  // extern "C"
