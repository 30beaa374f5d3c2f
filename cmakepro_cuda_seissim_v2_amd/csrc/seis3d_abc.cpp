// seis3d_abc.cpp - learned absorbing boundary, host side (no reference
// counterpart): attaching the 8-cell learned frame to a session, the GPU
// trainer (teacher-forced samples from wide-CPML runs, forward / backward /
// Adam on this library's kernels) and a gradient-check entry point.
//
// Frame: the 8 CPML cells of a pml = 8 context on every face except a free
// top.  While attached, the CPML tables are the identity (a = b = 0,
// 1/kappa = 1), so the step kernels give those cells the plain update; with
// parameters, every step then adds the MLP's correction to the 9 fields of
// the frame's outer 4 layers (the cells whose O(8) stencils reach the zero
// ghost cells; the inner 4 layers' plain update is already exact).
// Detaching restores the CPML tables: plain 8-cell CPML is the fallback.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/seis3d.h"
#include "seis3d_abc.hpp"
#include "seis3d_host.hpp"

namespace seis3d::host {

struct AbcFrame {
    int* cells = nullptr;
    unsigned char* codes = nullptr;
    int count = 0;
    void* image = nullptr;     // BF16 weights + FP32 biases; null: plain update only
    float* delta = nullptr;    // [count][9]
};

void release_abc(AbcFrame*& a) noexcept {
    if (!a) return;
    (void)hipFree(a->cells);
    (void)hipFree(a->codes);
    (void)hipFree(a->image);
    (void)hipFree(a->delta);
    delete a;
    a = nullptr;
}

namespace {

seis3d::abc::Frame frame_of(const seis3d_t* t) {
    const seis3d::DevStep P = t->base();
    seis3d::abc::Frame F{};
    F.g = P.g;
    for (int k = 0; k < 9; ++k) F.f[k] = P.p.f[k];
    F.inv_rho = P.p.inv_rho;
    F.lambda = P.p.lambda;
    F.mu = P.p.mu;
    return F;
}

seis3d::abc::FrameCells cells_of(const AbcFrame* a) { return {a->cells, a->codes, a->count}; }

// Corrected cells: non-ghost cells within 4 of an absorbing face (their stencils reach the ghost
// cells); the normal is the axis of the nearest face.
void enumerate(const seis3d_ctx& c, int px, std::vector<int>& cells, std::vector<unsigned char>& codes) {
    const int n[3] = {c.gx, c.gy, c.gz};
    for (int z = 4; z < c.gz - 4; ++z)
        for (int y = 4; y < c.gy - 4; ++y)
            for (int x = 4; x < c.gx - 4; ++x) {
                const int p[3] = {x, y, z};
                int best = 1 << 30;
                unsigned char code = 0;
                for (int a = 0; a < 3; ++a) {
                    const int lo = p[a] - 4, hi = n[a] - 5 - p[a];
                    if (!(a == 2 && c.free_surface) && lo < 8 && lo < best) { best = lo; code = static_cast<unsigned char>(a << 1); }
                    if (hi < 8 && hi < best) { best = hi; code = static_cast<unsigned char>((a << 1) | 1); }
                }
                if (best < 4) {
                    cells.push_back((z * c.gy + y) * px + x);
                    codes.push_back(code);
                }
            }
}

void upload_tables(seis3d_t* t, bool identity) {
    for (int a = 0; a < 3; ++a) {
        std::vector<float> v = make_coefficients(t->c, a, t->cpml_profile());
        if (identity)
            for (std::size_t k = 0; k < v.size(); k += 3) { v[k] = 0.0f; v[k + 1] = 0.0f; v[k + 2] = 1.0f; }
        check(hipMemcpyAsync(t->coef[a], v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice, t->stream),
              "upload frame tables");
    }
    check(hipStreamSynchronize(t->stream), "upload frame tables");
}

void upload_image(const float* params, void*& image, hipStream_t s) {
    float* p = nullptr;
    check(hipMalloc(&p, seis3d::abc::kParams * sizeof(float)), "allocate MLP parameters");
    try {
        check(hipMemcpyAsync(p, params, seis3d::abc::kParams * sizeof(float), hipMemcpyHostToDevice, s), "upload MLP");
        if (!image) check(hipMalloc(&image, seis3d::abc::kImageBytes), "allocate MLP image");
        check(seis3d::abc::launch_pack(p, image, s), "kernel launch");
        check(hipStreamSynchronize(s), "pack MLP");
    } catch (...) {
        (void)hipFree(p);
        throw;
    }
    (void)hipFree(p);
}

void check_params(const float* params) {
    for (int k = 0; k < seis3d::abc::kParams; ++k)
        if (!std::isfinite(params[k])) invalid("Seis3D: learned-boundary parameters must be finite.");
}

}  // namespace

void enqueue_abc(seis3d_t* t) {
    AbcFrame* a = t->abc;
    if (!a || !a->image || a->count == 0) return;
    const seis3d::abc::Frame F = frame_of(t);
    check(seis3d::abc::launch_correct(F, cells_of(a), a->image, a->delta, t->stream), "kernel launch");
    check(seis3d::abc::launch_apply(t->fields, cells_of(a), a->delta, t->stream), "kernel launch");
}

}  // namespace seis3d::host

using namespace seis3d::host;

struct seis3d_abc_trainer {
    hipStream_t stream = nullptr;
    int capacity = 0, count = 0;
    std::mt19937_64 rng;
    float *X = nullptr, *T = nullptr;
    float *p = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *partial = nullptr, *loss = nullptr;
    void* image = nullptr;

    void release() noexcept {
        if (stream) (void)hipStreamSynchronize(stream);
        for (float* q : {X, T, p, m, v, grad, partial, loss}) (void)hipFree(q);
        (void)hipFree(image);
        X = T = p = m = v = grad = partial = loss = nullptr;
        image = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

namespace {

constexpr int kChunks = 128;
constexpr float kMaxTarget = 4.0f;   // largest |normalised target| a training row may carry
constexpr int kMaxRetries = 6;       // diverged epochs redone at half rate before giving up

struct Scratch {
    std::vector<void*> ptrs;
    ~Scratch() { for (void* q : ptrs) (void)hipFree(q); }
    template <class T>
    T* get(std::size_t n) {
        void* q = nullptr;
        check(hipMalloc(&q, std::max<std::size_t>(1, n) * sizeof(T)), "allocate scratch");
        ptrs.push_back(q);
        return static_cast<T*>(q);
    }
};

// One forward + backward over rows [0, n) of X / T into grad (and per-sample errors into err).
void forward_backward(hipStream_t s, const void* image, const float* X, const float* T, int n, float denom, float* H1,
                      float* H2, float* Y, float* dY, float* dH2, float* dH1, float* err, float* partial, float* grad) {
    using namespace seis3d::abc;
    check(launch_mlp_forward(image, X, n, H1, H2, Y, s), "kernel launch");
    check(launch_mlp_backward_samples(image, H1, H2, Y, T, n, denom, dY, dH2, dH1, err, s), "kernel launch");
    check(launch_mlp_param_grads(X, H1, H2, dY, dH2, dH1, n, partial, kChunks, grad, s), "kernel launch");
}

}  // namespace

extern "C" {

int seis3d_abc_param_count(void) { return seis3d::abc::kParams; }

int seis3d_abc_attach(seis3d_t* t, const float* params) {
    return guarded([&] {
        require_unbatched(t, "attaching the learned boundary");
        const seis3d_ctx& c = t->c;
        if (c.reflecting || c.pml != 8)
            invalid("Seis3D: the learned boundary needs an 8-cell absorbing frame (pml = 8, no reflecting walls).");
        if (params) check_params(params);
        check(hipStreamSynchronize(t->stream), "attach learned boundary");
        if (!t->abc) {
            std::vector<int> cells;
            std::vector<unsigned char> codes;
            enumerate(c, t->plan.px, cells, codes);
            auto* a = new AbcFrame();
            try {
                a->count = int(cells.size());
                check(hipMalloc(&a->cells, cells.size() * sizeof(int)), "allocate frame");
                check(hipMalloc(&a->codes, codes.size()), "allocate frame");
                check(hipMalloc(&a->delta, cells.size() * seis3d::abc::kOut * sizeof(float)), "allocate frame");
                check(hipMemcpy(a->cells, cells.data(), cells.size() * sizeof(int), hipMemcpyHostToDevice), "upload frame");
                check(hipMemcpy(a->codes, codes.data(), codes.size(), hipMemcpyHostToDevice), "upload frame");
            } catch (...) {
                release_abc(a);
                throw;
            }
            t->abc = a;
        }
        if (params) upload_image(params, t->abc->image, t->stream);
        else {
            (void)hipFree(t->abc->image);
            t->abc->image = nullptr;
        }
        upload_tables(t, true);
    });
}

int seis3d_abc_detach(seis3d_t* t) {
    return guarded([&] {
        require_unbatched(t, "detaching the learned boundary");
        if (!t->abc) return;
        check(hipStreamSynchronize(t->stream), "detach learned boundary");
        release_abc(t->abc);
        upload_tables(t, false);
    });
}

int seis3d_abc_frame_cells(const seis3d_t* t, int* count) {
    return guarded([&] {
        require_session(t);
        if (count) *count = t->abc ? t->abc->count : 0;
    });
}

int seis3d_copy_wavefield_box(seis3d_t* dst, const seis3d_t* src, int ox, int oy, int oz) {
    return guarded([&] {
        require_session(dst);
        require_session(src);
        const seis3d_ctx &d = dst->c, &s = src->c;
        if (ox < 0 || oy < 0 || oz < 0 || d.gx - 4 + ox > s.gx || d.gy - 4 + oy > s.gy || d.gz - 4 + oz > s.gz)
            invalid("Seis3D: the destination box does not fit in the source grid.");
        check(hipStreamSynchronize(src->stream), "copy wavefield box");
        const seis3d::DevStep D = dst->base(), S = src->base();
        float* df[9];
        const float* sf[9];
        for (int k = 0; k < 9; ++k) { df[k] = D.p.f[k]; sf[k] = S.p.f[k]; }
        check(seis3d::abc::launch_copy_box(D.g, df, S.g, sf, ox, oy, oz, dst->stream), "kernel launch");
        check(hipStreamSynchronize(dst->stream), "copy wavefield box");
        dst->ghosts_valid = false;
    });
}

int seis3d_abc_trainer_create(int capacity, uint64_t seed, seis3d_abc_trainer_t** out) {
    return guarded([&] {
        if (capacity < 1 || capacity > (1 << 26) || !out) invalid("Seis3D: invalid trainer capacity.");
        auto* tr = new seis3d_abc_trainer();
        try {
            tr->capacity = capacity;
            tr->rng.seed(seed);
            check(hipStreamCreateWithFlags(&tr->stream, hipStreamNonBlocking), "create stream");
            const std::size_t cap = std::size_t(capacity);
            check(hipMalloc(&tr->X, cap * seis3d::abc::kIn * sizeof(float)), "allocate trainer");
            check(hipMalloc(&tr->T, cap * seis3d::abc::kOut * sizeof(float)), "allocate trainer");
            for (float** q : {&tr->p, &tr->m, &tr->v, &tr->grad})
                check(hipMalloc(q, seis3d::abc::kParams * sizeof(float)), "allocate trainer");
            check(hipMalloc(&tr->partial, std::size_t(kChunks) * seis3d::abc::kParams * sizeof(float)), "allocate trainer");
            check(hipMalloc(&tr->loss, 4096 * sizeof(float)), "allocate trainer");
            check(hipMalloc(&tr->image, seis3d::abc::kImageBytes), "allocate trainer");
        } catch (...) {
            tr->release();
            delete tr;
            throw;
        }
        *out = tr;
    });
}

void seis3d_abc_trainer_destroy(seis3d_abc_trainer_t* tr) {
    if (!tr) return;
    tr->release();
    delete tr;
}

int seis3d_abc_trainer_count(const seis3d_abc_trainer_t* tr) { return tr ? tr->count : 0; }

int seis3d_abc_trainer_add(seis3d_abc_trainer_t* tr, seis3d_t* thin, const seis3d_t* wide, int ox, int oy, int oz,
                           int max_cells, float min_rel_scale, int* added) {
    return guarded([&] {
        using namespace seis3d::abc;
        if (!tr) invalid("Seis3D: null trainer.");
        require_session(thin);
        require_session(wide);
        if (!thin->abc) fail(SEIS3D_ELIFECYCLE, "Seis3D: attach the learned boundary to the thin session first.");
        const seis3d_ctx &d = thin->c, &s = wide->c;
        if (ox < 0 || oy < 0 || oz < 0 || d.gx - 4 + ox > s.gx || d.gy - 4 + oy > s.gy || d.gz - 4 + oz > s.gz)
            invalid("Seis3D: the thin grid does not fit in the wide grid at this offset.");
        check(hipStreamSynchronize(thin->stream), "trainer add");
        check(hipStreamSynchronize(wide->stream), "trainer add");
        const int L = thin->abc->count;
        Scratch sc;
        float* Xall = sc.get<float>(std::size_t(L) * kIn);
        float* Sall = sc.get<float>(std::size_t(L));
        float* Ball = sc.get<float>(std::size_t(L) * kOut);
        float* Vall = sc.get<float>(std::size_t(L) * kOut);
        float* Tall = sc.get<float>(std::size_t(L) * kOut);
        float* Mall = sc.get<float>(std::size_t(L));
        const Frame Ft = frame_of(thin), Fw = frame_of(wide);
        const FrameCells C = cells_of(thin->abc);
        check(launch_features(Ft, C, Xall, Sall, Ball, tr->stream), "kernel launch");
        check(launch_face_values(Fw, Ft.g, C, ox, oy, oz, Vall, tr->stream), "kernel launch");
        check(launch_targets(Vall, Ball, Sall, L, Tall, Mall, tr->stream), "kernel launch");
        std::vector<float> hs(static_cast<std::size_t>(L)), hm(static_cast<std::size_t>(L));
        check(hipMemcpyAsync(hs.data(), Sall, hs.size() * sizeof(float), hipMemcpyDeviceToHost, tr->stream), "read scales");
        check(hipMemcpyAsync(hm.data(), Mall, hm.size() * sizeof(float), hipMemcpyDeviceToHost, tr->stream), "read scales");
        check(hipStreamSynchronize(tr->stream), "read scales");
        const float top = hs.empty() ? 0.0f : *std::max_element(hs.begin(), hs.end());
        // A one-step correction is a fraction of its window's largest value; rows
        // whose scale does not bound the target (a front entering an almost silent
        // window) are outliers that would dominate the squared loss.
        std::vector<int> cand;
        for (int k = 0; k < L; ++k)
            if (top > 0.0f && hs[std::size_t(k)] > min_rel_scale * top && hm[std::size_t(k)] <= kMaxTarget)
                cand.push_back(k);
        const int room = tr->capacity - tr->count;
        const int take = std::max(0, std::min({max_cells, room, int(cand.size())}));
        for (int k = 0; k < take; ++k) {   // partial Fisher-Yates
            std::uniform_int_distribution<int> u(k, int(cand.size()) - 1);
            std::swap(cand[std::size_t(k)], cand[std::size_t(u(tr->rng))]);
        }
        if (take > 0) {
            int* idx = sc.get<int>(std::size_t(take));
            check(hipMemcpyAsync(idx, cand.data(), std::size_t(take) * sizeof(int), hipMemcpyHostToDevice, tr->stream),
                  "upload sample rows");
            check(launch_gather_rows(Xall, idx, take, kIn, tr->X, tr->count, tr->stream), "kernel launch");
            check(launch_gather_rows(Tall, idx, take, kOut, tr->T, tr->count, tr->stream), "kernel launch");
            check(hipStreamSynchronize(tr->stream), "trainer add");
        }
        tr->count += take;
        if (added) *added = take;
    });
}

int seis3d_abc_train(seis3d_abc_trainer_t* tr, float* params, int epochs, int batch, float lr, float* epoch_loss) {
    return guarded([&] {
        using namespace seis3d::abc;
        if (!tr || !params || epochs < 1 || batch < 16 || batch > (1 << 20) || !(lr > 0.0f))
            invalid("Seis3D: invalid training request.");
        if (tr->count < 16) invalid("Seis3D: the trainer holds too few samples.");
        check_params(params);
        const int n = tr->count;
        const int batches = (n + batch - 1) / batch;
        if (batches > 4096) invalid("Seis3D: more than 4096 batches per epoch; raise the batch size.");
        Scratch sc;
        float* Xs = sc.get<float>(std::size_t(n) * kIn);
        float* Ts = sc.get<float>(std::size_t(n) * kOut);
        float* H1 = sc.get<float>(std::size_t(batch) * kHidden);
        float* H2 = sc.get<float>(std::size_t(batch) * kHidden);
        float* Y = sc.get<float>(std::size_t(batch) * kOut);
        float* dY = sc.get<float>(std::size_t(batch) * kOut);
        float* dH2 = sc.get<float>(std::size_t(batch) * kHidden);
        float* dH1 = sc.get<float>(std::size_t(batch) * kHidden);
        float* err = sc.get<float>(std::size_t(batch));
        float* sum_partial = sc.get<float>(256);
        int* perm = sc.get<int>(std::size_t(n));
        std::vector<int> order(static_cast<std::size_t>(n));
        std::iota(order.begin(), order.end(), 0);
        std::shuffle(order.begin(), order.end(), tr->rng);
        hipStream_t s = tr->stream;
        check(hipMemcpyAsync(perm, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice, s), "upload order");
        check(launch_gather_rows(tr->X, perm, n, kIn, Xs, 0, s), "kernel launch");
        check(launch_gather_rows(tr->T, perm, n, kOut, Ts, 0, s), "kernel launch");
        check(hipMemcpyAsync(tr->p, params, kParams * sizeof(float), hipMemcpyHostToDevice, s), "upload parameters");
        check(hipMemsetAsync(tr->m, 0, kParams * sizeof(float), s), "clear moments");
        check(hipMemsetAsync(tr->v, 0, kParams * sizeof(float), s), "clear moments");
        check(launch_pack(tr->p, tr->image, s), "kernel launch");
        int step = 0;
        std::vector<float> losses(static_cast<std::size_t>(batches));
        std::vector<float> host_p(static_cast<std::size_t>(kParams));
        float* backup = sc.get<float>(3 * std::size_t(kParams));
        float damping = 1.0f;
        int retries = 0;
        for (int e = 0; e < epochs; ++e) {
            // Cosine decay from lr to lr / 20 (halved after each diverged epoch).
            const float rate = damping * lr *
                               (0.05f + 0.95f * 0.5f * (1.0f + std::cos(3.14159265f * float(e) / float(epochs))));
            const int step0 = step;
            check(hipMemcpyAsync(backup, tr->p, kParams * sizeof(float), hipMemcpyDeviceToDevice, s), "save state");
            check(hipMemcpyAsync(backup + kParams, tr->m, kParams * sizeof(float), hipMemcpyDeviceToDevice, s),
                  "save state");
            check(hipMemcpyAsync(backup + 2 * kParams, tr->v, kParams * sizeof(float), hipMemcpyDeviceToDevice, s),
                  "save state");
            for (int b = 0; b < batches; ++b) {
                const int first = b * batch, rows = std::min(batch, n - first);
                forward_backward(s, tr->image, Xs + std::size_t(first) * kIn, Ts + std::size_t(first) * kOut, rows,
                                 float(rows) * kOut, H1, H2, Y, dY, dH2, dH1, err, tr->partial, tr->grad);
                check(launch_sum(err, rows, sum_partial, tr->loss + b, s), "kernel launch");
                check(launch_adam(tr->p, tr->m, tr->v, tr->grad, rate, 0.9f, 0.999f, 1e-8f, ++step, tr->image, s),
                      "kernel launch");
            }
            check(hipMemcpyAsync(losses.data(), tr->loss, losses.size() * sizeof(float), hipMemcpyDeviceToHost, s),
                  "read losses");
            check(hipMemcpyAsync(host_p.data(), tr->p, kParams * sizeof(float), hipMemcpyDeviceToHost, s),
                  "read parameters");
            check(hipStreamSynchronize(s), "train");
            double total = 0;
            for (float l : losses) total += double(l);
            bool finite = std::isfinite(total);
            for (float v : host_p) finite = finite && std::isfinite(v);
            if (!finite) {
                // Diverged: restore the epoch's starting state and redo it at half the rate.
                if (++retries > kMaxRetries)
                    fail(SEIS3D_ENONFINITE, "Seis3D: learned-boundary training diverged; lower the learning rate.");
                check(hipMemcpyAsync(tr->p, backup, kParams * sizeof(float), hipMemcpyDeviceToDevice, s), "restore state");
                check(hipMemcpyAsync(tr->m, backup + kParams, kParams * sizeof(float), hipMemcpyDeviceToDevice, s),
                      "restore state");
                check(hipMemcpyAsync(tr->v, backup + 2 * kParams, kParams * sizeof(float), hipMemcpyDeviceToDevice, s),
                      "restore state");
                check(launch_pack(tr->p, tr->image, s), "kernel launch");
                step = step0;
                damping *= 0.5f;
                --e;
                continue;
            }
            if (epoch_loss) epoch_loss[e] = float(total / (double(n) * kOut));
        }
        check(hipMemcpyAsync(params, tr->p, kParams * sizeof(float), hipMemcpyDeviceToHost, s), "read parameters");
        check(hipStreamSynchronize(s), "train");
    });
}

int seis3d_abc_evaluate(seis3d_abc_trainer_t* tr, const float* params, double* loss) {
    return guarded([&] {
        using namespace seis3d::abc;
        if (!tr || !params || !loss || tr->count < 1) invalid("Seis3D: invalid evaluation request.");
        check_params(params);
        const int n = tr->count;
        Scratch sc;
        float* H1 = sc.get<float>(std::size_t(n) * kHidden);
        float* H2 = sc.get<float>(std::size_t(n) * kHidden);
        float* Y = sc.get<float>(std::size_t(n) * kOut);
        float* dY = sc.get<float>(std::size_t(n) * kOut);
        float* dH2 = sc.get<float>(std::size_t(n) * kHidden);
        float* dH1 = sc.get<float>(std::size_t(n) * kHidden);
        float* err = sc.get<float>(std::size_t(n));
        float* sum_partial = sc.get<float>(256);
        hipStream_t s = tr->stream;
        check(hipMemcpyAsync(tr->p, params, kParams * sizeof(float), hipMemcpyHostToDevice, s), "upload parameters");
        check(launch_pack(tr->p, tr->image, s), "kernel launch");
        forward_backward(s, tr->image, tr->X, tr->T, n, float(n) * kOut, H1, H2, Y, dY, dH2, dH1, err, tr->partial,
                         tr->grad);
        check(launch_sum(err, n, sum_partial, tr->loss, s), "kernel launch");
        float total = 0;
        check(hipMemcpyAsync(&total, tr->loss, sizeof(float), hipMemcpyDeviceToHost, s), "read loss");
        check(hipStreamSynchronize(s), "evaluate");
        *loss = double(total) / (double(n) * kOut);
    });
}

int seis3d_abc_mlp(const float* params, const float* X, const float* T, int n, float* Y, float* H1, float* H2,
                   float* grad, double* loss) {
    return guarded([&] {
        using namespace seis3d::abc;
        if (!params || !X || !T || n < 1 || n > (1 << 22)) invalid("Seis3D: invalid MLP request.");
        hipStream_t s = nullptr;
        check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "create stream");
        struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); } } guard{s};
        Scratch sc;
        float* p = sc.get<float>(kParams);
        void* image = sc.get<unsigned char>(kImageBytes);
        float* dX = sc.get<float>(std::size_t(n) * kIn);
        float* dT = sc.get<float>(std::size_t(n) * kOut);
        float* dH1v = sc.get<float>(std::size_t(n) * kHidden);
        float* dH2v = sc.get<float>(std::size_t(n) * kHidden);
        float* dYv = sc.get<float>(std::size_t(n) * kOut);
        float* gY = sc.get<float>(std::size_t(n) * kOut);
        float* gH2 = sc.get<float>(std::size_t(n) * kHidden);
        float* gH1 = sc.get<float>(std::size_t(n) * kHidden);
        float* err = sc.get<float>(std::size_t(n));
        float* partial = sc.get<float>(std::size_t(kChunks) * kParams);
        float* g = sc.get<float>(kParams);
        float* sum_partial = sc.get<float>(256);
        float* total = sc.get<float>(1);
        check(hipMemcpyAsync(p, params, kParams * sizeof(float), hipMemcpyHostToDevice, s), "upload");
        check(hipMemcpyAsync(dX, X, std::size_t(n) * kIn * sizeof(float), hipMemcpyHostToDevice, s), "upload");
        check(hipMemcpyAsync(dT, T, std::size_t(n) * kOut * sizeof(float), hipMemcpyHostToDevice, s), "upload");
        check(launch_pack(p, image, s), "kernel launch");
        forward_backward(s, image, dX, dT, n, float(n) * kOut, dH1v, dH2v, dYv, gY, gH2, gH1, err, partial, g);
        check(launch_sum(err, n, sum_partial, total, s), "kernel launch");
        if (Y) check(hipMemcpyAsync(Y, dYv, std::size_t(n) * kOut * sizeof(float), hipMemcpyDeviceToHost, s), "read");
        if (H1) check(hipMemcpyAsync(H1, dH1v, std::size_t(n) * kHidden * sizeof(float), hipMemcpyDeviceToHost, s), "read");
        if (H2) check(hipMemcpyAsync(H2, dH2v, std::size_t(n) * kHidden * sizeof(float), hipMemcpyDeviceToHost, s), "read");
        if (grad) check(hipMemcpyAsync(grad, g, kParams * sizeof(float), hipMemcpyDeviceToHost, s), "read");
        float tl = 0;
        check(hipMemcpyAsync(&tl, total, sizeof(float), hipMemcpyDeviceToHost, s), "read");
        check(hipStreamSynchronize(s), "MLP");
        if (loss) *loss = double(tl) / (double(n) * kOut);
    });
}

}  // extern "C"
