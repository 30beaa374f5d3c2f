// seis3d_fwi.cpp - FWI gradient driver (no reference counterpart): forward
// modelling with checkpoints, segment recompute and the discrete-adjoint
// sweep, for one or many shots that share the session's model.
//
// Memory plan per shot (all in HBM, one allocation carved into parts):
//  * live forward state: 9 wavefields + CPML memories ("state");
//  * checkpoints: the state at steps K, 2K, ... before the last segment
//    (step 0 is the zero state and the last segment is stored while the
//    first forward pass runs, so neither is kept as a checkpoint);
//  * a slot store of K slots for the segment being reversed: slot m holds the
//    stresses before step s0+m and the velocities and CPML memories after it,
//    exactly what the adjoint of that step reads;
//  * adjoint state: 9 adjoint wavefields, adjoint CPML memories, 9 filtered
//    derivative temporaries, 8 gradient sums, 3 surface scratch planes.
// Shots run on their own streams, enqueued in lockstep so that they progress
// together; as many run at once as fit in device memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/seis3d.h"
#include "seis3d_adjoint.hpp"
#include "seis3d_host.hpp"

using namespace seis3d::host;

namespace {


struct StateView {
    float* f[9];
    float* psi[3];
};

// Floats of each part of a state, each part padded to 64 floats (256 bytes).
struct StateLayout {
    std::size_t cells = 0, plane = 0, psi[3]{};
    static std::size_t pad(std::size_t n) { return align_up(n, 64); }
    std::size_t floats() const { return 9 * pad(cells) + pad(psi[0]) + pad(psi[1]) + pad(psi[2]); }
    StateView view(float* base) const {
        StateView v{};
        for (int k = 0; k < 9; ++k) { v.f[k] = base; base += pad(cells); }
        for (int a = 0; a < 3; ++a) { v.psi[a] = base; base += pad(psi[a]); }
        return v;
    }
};

StateLayout layout_for(const Plan& p) {
    StateLayout L;
    L.cells = p.cells;
    L.plane = p.plane;
    for (int a = 0; a < 3; ++a) L.psi[a] = p.slab[a] * 6;
    return L;
}

struct FwiPlan {
    int64_t steps = 0;
    int stride = 1, interval = 1, segments = 1, samples = 0;
    std::size_t state_floats = 0, adjoint_floats = 0, shot_bytes = 0;
};

std::size_t adjoint_floats(const StateLayout& L) {
    // 9 adjoint fields + adjoint psi + 9 temporaries + 8 gradient sums + 3 surface planes
    return 26 * StateLayout::pad(L.cells) + 3 * StateLayout::pad(L.plane) + StateLayout::pad(L.psi[0]) +
           StateLayout::pad(L.psi[1]) + StateLayout::pad(L.psi[2]);
}

FwiPlan fwi_plan(const seis3d_ctx& c, int64_t steps, int stride, int interval, int receivers) {
    if (steps < 1 || steps > 1000000) invalid("Seis3D: FWI step count must be 1..1000000.");
    if (stride < 1 || stride > steps) invalid("Seis3D: FWI sample stride must be 1..n_steps.");
    if (interval <= 0) interval = int(std::ceil(std::sqrt(double(steps))));
    if (interval > steps) interval = int(steps);
    if (receivers < 1 || receivers > kMaxReceivers) invalid("Seis3D: FWI shots need 1..300 receivers.");
    FwiPlan f;
    f.steps = steps;
    f.stride = stride;
    f.interval = interval;
    f.segments = int((steps + interval - 1) / interval);
    f.samples = int(steps / stride);
    const StateLayout L = layout_for(plan_for(c));
    f.state_floats = L.floats();
    f.adjoint_floats = adjoint_floats(L);
    const std::size_t records = 3 * StateLayout::pad(std::size_t(f.samples) * receivers * 3);
    const std::size_t checkpoints = std::size_t(std::max(0, f.segments - 2));
    f.shot_bytes = ((1 + checkpoints + std::size_t(interval)) * f.state_floats + f.adjoint_floats + records) *
                       sizeof(float) +
                   StateLayout::pad(std::size_t(receivers) * 3) * sizeof(int) + 256;
    return f;
}

struct ShotWork {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    char* mem = nullptr;
    unsigned int* flags = nullptr;
    seis3d_source src{};
    int nrec = 0;
    int* rec = nullptr;
    float* gather = nullptr;
    float* observed = nullptr;
    float* residual = nullptr;
    StateView live{};
    std::vector<StateView> checkpoints;   // index j-1 holds the state at step j K
    std::vector<StateView> slots;
    float* a[9]{};
    float* apsi[3]{};
    float* t[9]{};
    float* acc[8]{};                      // g_lam, g_dev, g_shear[3], g_buoy[3]
    float* surf[3]{};                     // free-surface scratch planes

    void release() noexcept {
        if (stream) (void)hipStreamSynchronize(stream);
        (void)hipFree(mem);
        (void)hipFree(flags);
        if (done) (void)hipEventDestroy(done);
        if (stream) (void)hipStreamDestroy(stream);
        mem = nullptr; flags = nullptr; done = nullptr; stream = nullptr;
    }
};

struct Carver {
    char* p;
    float* floats(std::size_t n) {
        float* r = reinterpret_cast<float*>(p);
        p += StateLayout::pad(n) * sizeof(float);
        return r;
    }
};

void allocate_shot(ShotWork& w, const FwiPlan& f, const StateLayout& L, int nrec) {
    check(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking), "create FWI stream");
    check(hipEventCreateWithFlags(&w.done, hipEventDisableTiming), "create FWI event");
    check(hipMalloc(&w.mem, f.shot_bytes), "allocate FWI workspace");
    check(hipMalloc(&w.flags, sizeof(unsigned int)), "allocate FWI flags");
    Carver cv{w.mem};
    w.live = L.view(cv.floats(f.state_floats));
    for (int j = 0; j < std::max(0, f.segments - 2); ++j) w.checkpoints.push_back(L.view(cv.floats(f.state_floats)));
    for (int m = 0; m < f.interval; ++m) w.slots.push_back(L.view(cv.floats(f.state_floats)));
    for (auto& p : w.a) p = cv.floats(L.cells);
    for (int a = 0; a < 3; ++a) w.apsi[a] = cv.floats(L.psi[a]);
    for (auto& p : w.t) p = cv.floats(L.cells);
    for (auto& p : w.acc) p = cv.floats(L.cells);
    for (auto& p : w.surf) p = cv.floats(L.plane);
    const std::size_t record = std::size_t(f.samples) * nrec * 3;
    w.gather = cv.floats(record);
    w.observed = cv.floats(record);
    w.residual = cv.floats(record);
    w.rec = reinterpret_cast<int*>(cv.p);
    w.nrec = nrec;
}

// Enqueues fills/copies in batches of at most 16 arrays.
struct Batcher {
    hipStream_t s;
    int op;   // 0 copy, 1 zero, 2 add
    seis3d::ArrayBatch b{};
    void add(const float* src, float* dst, std::size_t n) {
        if (n == 0) return;
        b.src[b.count] = src;
        b.dst[b.count] = dst;
        b.n[b.count] = static_cast<long long>(n);
        if (++b.count == 16) flush();
    }
    void flush() {
        if (!b.count) return;
        const hipError_t e = op == 0 ? seis3d::launch_copy_arrays(b, s)
                           : op == 1 ? seis3d::launch_zero_arrays(b, s)
                                     : seis3d::launch_add_arrays(b, s);
        check(e, "array kernel launch");
        b = {};
    }
};

void zero_state(hipStream_t s, const StateView& v, const StateLayout& L) {
    Batcher z{s, 1};
    for (auto* p : v.f) z.add(nullptr, p, L.cells);
    for (int a = 0; a < 3; ++a) z.add(nullptr, v.psi[a], L.psi[a]);
    z.flush();
}

struct Job {
    seis3d_t* t;
    const seis3d_ctx& c;
    FwiPlan f;
    StateLayout L;
    seis3d::DevStep base;

    seis3d::DevStep forward_operands(const ShotWork& w) const {
        seis3d::DevStep P = base;
        for (int k = 0; k < 9; ++k) P.p.f[k] = P.p.out[k] = w.live.f[k];
        for (int a = 0; a < 3; ++a) P.p.psi[a] = P.p.psi_out[a] = w.live.psi[a];
        P.p.flags = w.flags;
        return P;
    }

    // One step n (runStep3D, cuda/Cuda_Seis3D.cu:602-624).  The stress ghosts and
    // walls that open a step are already in place: the state starts at zero and
    // every step ends with them.
    void forward_step(ShotWork& w, int64_t n) const {
        seis3d::DevStep P = forward_operands(w);
        seis3d::DevStep V = P;
        if (w.src.kind == SEIS3D_FORCE) V.src = make_source(c, w.src, double(n) * c.dt);
        check(seis3d::launch_velocity(V, t->kernels, w.stream), "kernel launch");
        if (c.free_surface) check(seis3d::launch_velocity_ghost(P, c.spacing, w.stream), "kernel launch");
        if (c.reflecting) check(seis3d::launch_reflecting_walls(P, true, w.stream), "kernel launch");
        seis3d::DevStep S = P;
        if (w.src.kind == SEIS3D_EXPLOSION) S.src = make_source(c, w.src, (double(n) + 0.5) * c.dt);
        check(seis3d::launch_stress(S, t->kernels, w.stream), "kernel launch");
        // The LDS-DMA stress kernel writes the stress ghost planes itself.
        if (c.free_surface && t->kernels.family != 4)
            check(seis3d::launch_stress_ghost(P, w.stream), "kernel launch");
        if (c.reflecting) check(seis3d::launch_reflecting_walls(P, false, w.stream), "kernel launch");
    }

    // Step n = s0 + m of a recomputed segment, out of place, straight into the slots:
    // from `prev` (the segment's start state for m = 0, else slot m - 1) into slot m,
    // which holds the velocities and CPML memories after step n and, already, the
    // stresses before it.  The stresses after step n go to slot m + 1; after the
    // segment's last step nobody reads them and they go to the live state.  The
    // kernels and their arithmetic are those of forward_step, without the copies.
    // LDS-DMA kernels (family 4) only: the other families update in place.
    void recompute_step(ShotWork& w, int64_t n, int64_t s0, int64_t s1, const StateView& start) const {
        const std::size_t m = std::size_t(n - s0);
        const StateView& prev = m == 0 ? start : w.slots[m - 1];
        const StateView& cur = w.slots[m];
        const StateView& next = n + 1 < s1 ? w.slots[m + 1] : w.live;
        seis3d::DevStep V = base;
        V.p.flags = w.flags;
        for (int k = 0; k < 3; ++k) { V.p.f[k] = prev.f[k]; V.p.out[k] = cur.f[k]; }
        for (int k = 3; k < 9; ++k) V.p.f[k] = V.p.out[k] = cur.f[k];
        for (int a = 0; a < 3; ++a) { V.p.psi[a] = prev.psi[a]; V.p.psi_out[a] = cur.psi[a]; }
        seis3d::DevStep S = V;
        if (w.src.kind == SEIS3D_FORCE) V.src = make_source(c, w.src, double(n) * c.dt);
        check(seis3d::launch_velocity(V, t->kernels, w.stream), "kernel launch");
        for (int k = 0; k < 3; ++k) S.p.f[k] = S.p.out[k] = cur.f[k];
        for (int k = 3; k < 9; ++k) S.p.out[k] = next.f[k];
        // Velocity ghosts and walls in place in slot m; the stress kernel writes the
        // stress ghost planes into the arrays it writes, and the stress walls follow there.
        if (c.free_surface) check(seis3d::launch_velocity_ghost(S, c.spacing, w.stream), "kernel launch");
        if (c.reflecting) check(seis3d::launch_reflecting_walls(S, true, w.stream), "kernel launch");
        if (w.src.kind == SEIS3D_EXPLOSION) S.src = make_source(c, w.src, (double(n) + 0.5) * c.dt);
        check(seis3d::launch_stress(S, t->kernels, w.stream), "kernel launch");
        if (c.reflecting) {
            seis3d::DevStep W = S;
            for (int k = 3; k < 9; ++k) W.p.f[k] = W.p.out[k] = next.f[k];
            check(seis3d::launch_reflecting_walls(W, false, w.stream), "kernel launch");
        }
    }

    // After step n inside the segment starting at s0: velocities and psi into
    // slot n - s0, stresses into the σ part of the next slot.
    void store_slot(ShotWork& w, int64_t n, int64_t s0, int64_t s1) const {
        Batcher cp{w.stream, 0};
        const StateView& slot = w.slots[std::size_t(n - s0)];
        for (int k = 0; k < 3; ++k) cp.add(w.live.f[k], slot.f[k], L.cells);
        for (int a = 0; a < 3; ++a) cp.add(w.live.psi[a], slot.psi[a], L.psi[a]);
        if (n + 1 < s1) {
            const StateView& next = w.slots[std::size_t(n + 1 - s0)];
            for (int k = 3; k < 9; ++k) cp.add(w.live.f[k], next.f[k], L.cells);
        }
        cp.flush();
    }

    void store_first_stresses(ShotWork& w, const StateView& from) const {
        Batcher cp{w.stream, 0};
        for (int k = 3; k < 9; ++k) cp.add(from.f[k], w.slots[0].f[k], L.cells);
        cp.flush();
    }

    void copy_state(hipStream_t s, const StateView& from, const StateView& to) const {
        Batcher cp{s, 0};
        for (int k = 0; k < 9; ++k) cp.add(from.f[k], to.f[k], L.cells);
        for (int a = 0; a < 3; ++a) cp.add(from.psi[a], to.psi[a], L.psi[a]);
        cp.flush();
    }

    seis3d::AdjStep adjoint_operands(ShotWork& w, const StateView& slot, int64_t n) const {
        seis3d::AdjStep A{};
        A.g = base.g;
        for (int k = 0; k < 3; ++k) A.fv[k] = slot.f[k];
        for (int k = 0; k < 6; ++k) A.fs[k] = slot.f[3 + k];
        for (int a = 0; a < 3; ++a) { A.psi[a] = slot.psi[a]; A.coef[a] = base.p.coef[a]; A.apsi[a] = w.apsi[a]; }
        A.inv_rho = base.p.inv_rho; A.lambda = base.p.lambda; A.mu = base.p.mu; A.inv_mu = base.p.inv_mu;
        for (int k = 0; k < 9; ++k) { A.a[k] = w.a[k]; A.t[k] = w.t[k]; }
        A.g_lam = w.acc[0]; A.g_dev = w.acc[1];
        for (int k = 0; k < 3; ++k) { A.g_shear[k] = w.acc[2 + k]; A.g_buoy[k] = w.acc[5 + k]; }
        for (int k = 0; k < 3; ++k) A.surf[k] = w.surf[k];
        A.dt = base.dt;
        A.invh = base.invh;
        A.spacing = c.spacing;
        A.reflecting = c.reflecting;
        if (w.src.kind == SEIS3D_FORCE) A.src = make_source(c, w.src, double(n) * c.dt);
        return A;
    }

    void adjoint_step(ShotWork& w, int64_t n, int64_t s0) const {
        const StateView& slot = w.slots[std::size_t(n - s0)];
        seis3d::AdjStep A = adjoint_operands(w, slot, n);
        if ((n + 1) % f.stride == 0 && (n + 1) / f.stride <= f.samples) {
            const float* r = w.residual + std::size_t((n + 1) / f.stride - 1) * w.nrec * 3;
            check(seis3d::launch_adjoint_inject(A, w.rec, w.nrec, r, w.stream), "kernel launch");
        }
        check(seis3d::launch_adjoint_stress(A, w.stream), "kernel launch");
        check(seis3d::launch_adjoint_velocity(A, w.stream), "kernel launch");
    }

    // Whole gradient of a wave of shots, summed into `total` (8 accumulators).
    void run_wave(std::vector<ShotWork*>& shots, float* const total[8]) const {
        const int64_t K = f.interval, N = f.steps;
        const int J = f.segments;
        const int64_t last0 = int64_t(J - 1) * K;
        for (auto* w : shots) {
            check(hipMemsetAsync(w->flags, 0, sizeof(unsigned int), w->stream), "clear FWI flags");
            zero_state(w->stream, w->live, L);
            // Recomputed steps write no ghost cells but the free surface's and the walls',
            // so every slot must hold zero ghosts elsewhere.  The forward pass copies whole states into the
            // slots of the last segment; only the slots past a short last segment start
            // unwritten.
            for (std::size_t m = std::size_t(N - last0); m < w->slots.size(); ++m)
                zero_state(w->stream, w->slots[m], L);
            Batcher z{w->stream, 1};
            for (auto* p : w->a) z.add(nullptr, p, L.cells);
            for (int a = 0; a < 3; ++a) z.add(nullptr, w->apsi[a], L.psi[a]);
            for (auto* p : w->t) z.add(nullptr, p, L.cells);
            for (auto* p : w->acc) z.add(nullptr, p, L.cells);
            for (auto* p : w->surf) z.add(nullptr, p, L.plane);
            z.flush();
        }
        // Forward pass: receivers, checkpoints, and the last segment's slots.
        for (int64_t n = 0; n < N; ++n) {
            for (auto* w : shots) {
                if (n % K == 0 && n > 0 && n < last0) copy_state(w->stream, w->live, w->checkpoints[std::size_t(n / K - 1)]);
                if (n == last0) store_first_stresses(*w, w->live);
                forward_step(*w, n);
                if (n >= last0) store_slot(*w, n, last0, N);
                if ((n + 1) % f.stride == 0 && (n + 1) / f.stride <= f.samples) {
                    float* dst = w->gather + std::size_t((n + 1) / f.stride - 1) * w->nrec * 3;
                    check(seis3d::launch_receivers(forward_operands(*w), w->rec, w->nrec, dst, w->stream),
                          "kernel launch");
                }
            }
        }
        for (auto* w : shots)
            check(seis3d::launch_residual(w->gather, w->observed, w->residual,
                                          static_cast<long long>(f.samples) * w->nrec * 3, w->stream),
                  "kernel launch");
        // Backward: segments in reverse, recomputing every one but the last.
        for (int j = J - 1; j >= 0; --j) {
            const int64_t s0 = int64_t(j) * K, s1 = std::min(N, s0 + K);
            if (j != J - 1 && t->kernels.family == 4) {
                for (auto* w : shots) {
                    if (j == 0) zero_state(w->stream, w->live, L);
                    store_first_stresses(*w, j == 0 ? w->live : w->checkpoints[std::size_t(j - 1)]);
                }
                for (int64_t n = s0; n < s1; ++n)
                    for (auto* w : shots)
                        recompute_step(*w, n, s0, s1, j == 0 ? w->live : w->checkpoints[std::size_t(j - 1)]);
            } else if (j != J - 1) {
                // Kernel families without an out-of-place form: step the live state and copy.
                for (auto* w : shots) {
                    if (j == 0) zero_state(w->stream, w->live, L);
                    else copy_state(w->stream, w->checkpoints[std::size_t(j - 1)], w->live);
                    store_first_stresses(*w, w->live);
                }
                for (int64_t n = s0; n < s1; ++n)
                    for (auto* w : shots) {
                        forward_step(*w, n);
                        store_slot(*w, n, s0, s1);
                    }
            }
            for (int64_t n = s1 - 1; n >= s0; --n)
                for (auto* w : shots) adjoint_step(*w, n, s0);
        }
        // Sum each shot's gradient into the totals on the first shot's stream.
        hipStream_t s = shots[0]->stream;
        for (std::size_t k = 1; k < shots.size(); ++k) {
            check(hipEventRecord(shots[k]->done, shots[k]->stream), "record FWI event");
            check(hipStreamWaitEvent(s, shots[k]->done, 0), "join FWI streams");
        }
        // One launch per shot: launches on one stream are ordered, while the
        // arrays inside one launch run concurrently and must not share a target.
        for (auto* w : shots) {
            Batcher add{s, 2};
            for (int k = 0; k < 8; ++k) add.add(w->acc[k], total[k], L.cells);
            add.flush();
        }
        check(hipStreamSynchronize(s), "FWI gradient");
        for (auto* w : shots) check(hipStreamSynchronize(w->stream), "FWI gradient");
    }
};

}  // namespace

extern "C" {

int seis3d_fwi_plan(const seis3d_ctx* ctx, int64_t n_steps, int stride, int checkpoint_interval, int receivers,
                    uint64_t* shot_bytes, int* interval) {
    return guarded([&] {
        if (!ctx) invalid("Seis3D: null context.");
        seis3d_ctx c = *ctx;
        if (seis3d_validate_ctx(&c) != SEIS3D_OK) fail(SEIS3D_EINVAL, g_error);
        const FwiPlan f = fwi_plan(c, n_steps, stride, checkpoint_interval, receivers);
        if (shot_bytes) *shot_bytes = f.shot_bytes;
        if (interval) *interval = f.interval;
    });
}

int seis3d_fwi_gradient(seis3d_t* t, const seis3d_shot* shots, int count, int64_t n_steps, int stride,
                        int checkpoint_interval, double* misfit, float* g_vp, float* g_vs, float* g_rho) {
    return guarded([&] {
        require_session(t);
        const seis3d_ctx& c = t->c;
        if (!shots || count < 1 || count > 1024 || !g_vp || !g_vs || !g_rho)
            invalid("Seis3D: invalid FWI gradient request.");
        if (t->recording) fail(SEIS3D_ELIFECYCLE, "Seis3D: stop the gather before computing a gradient.");
        // The gradient is the adjoint of the plain CPML step; a learned frame would make it
        // the gradient of a different forward operator than seis3d_run's.
        if (t->abc) fail(SEIS3D_ELIFECYCLE, "Seis3D: detach the learned boundary before computing a gradient.");
        int most = 0;
        for (int s = 0; s < count; ++s) {
            check_source(c, shots[s].source);
            const int n = shots[s].receiver_count;
            if (n < 1 || n > kMaxReceivers || !shots[s].receivers || !shots[s].observed)
                invalid("Seis3D: every FWI shot needs 1..300 receivers and observed data.");
            for (int r = 0; r < n; ++r) {
                const auto p = shots[s].receivers[r];
                if (p.x < 0 || p.x >= c.nx || p.y < 0 || p.y >= c.ny || p.z < 0 || p.z >= c.nz)
                    invalid("Seis3D: receiver outside the physical domain.");
            }
            most = std::max(most, n);
        }
        const FwiPlan f = fwi_plan(c, n_steps, stride, checkpoint_interval, most);
        const StateLayout L = layout_for(t->plan);

        // The model (and 1/mu) must be complete before the shot streams start.
        seis3d::DevStep base = t->base();
        if (t->inv_mu_dirty) {
            check(seis3d::launch_reciprocal_mu(base, t->stream), "kernel launch");
            t->inv_mu_dirty = false;
        }
        check(hipStreamSynchronize(t->stream), "FWI gradient");

        std::size_t free_bytes = 0, total_bytes = 0;
        check(hipMemGetInfo(&free_bytes, &total_bytes), "query device memory");
        const std::size_t totals_bytes = 8 * StateLayout::pad(L.cells) * sizeof(float);
        if (free_bytes < totals_bytes + f.shot_bytes)
            invalid("Seis3D: the FWI checkpoint plan does not fit in device memory.");
        const std::size_t usable = std::size_t(double(free_bytes - totals_bytes) * 0.94);
        const int fit = int(std::max<std::size_t>(1, std::min<std::size_t>(std::size_t(count), usable / f.shot_bytes)));
        // Equal waves (8 shots with room for 7 run as 4 + 4, not 7 + 1).
        const int waves = (count + fit - 1) / fit;
        const int wave = (count + waves - 1) / waves;

        float* totals_mem = nullptr;
        std::vector<ShotWork> work(static_cast<std::size_t>(wave));
        double sum = 0;
        try {
            check(hipMalloc(&totals_mem, totals_bytes), "allocate FWI gradient");
            float* total[8];
            for (int k = 0; k < 8; ++k) total[k] = totals_mem + k * StateLayout::pad(L.cells);
            {
                Batcher z{t->stream, 1};
                for (auto* p : total) z.add(nullptr, p, L.cells);
                z.flush();
                check(hipStreamSynchronize(t->stream), "FWI gradient");
            }
            for (auto& w : work) allocate_shot(w, f, L, most);
            Job job{t, c, f, L, base};
            std::vector<float> host;
            for (int first = 0; first < count; first += wave) {
                const int n_here = std::min(wave, count - first);
                std::vector<ShotWork*> active;
                for (int k = 0; k < n_here; ++k) {
                    ShotWork& w = work[std::size_t(k)];
                    const seis3d_shot& sh = shots[first + k];
                    w.src = sh.source;
                    w.nrec = sh.receiver_count;
                    std::vector<int> padded(std::size_t(w.nrec) * 3);
                    for (int r = 0; r < w.nrec; ++r) {
                        padded[3 * r] = sh.receivers[r].x + c.ox;
                        padded[3 * r + 1] = sh.receivers[r].y + c.oy;
                        padded[3 * r + 2] = sh.receivers[r].z + c.oz;
                    }
                    const std::size_t record = std::size_t(f.samples) * w.nrec * 3;
                    check(hipMemcpyAsync(w.rec, padded.data(), padded.size() * sizeof(int), hipMemcpyHostToDevice,
                                         w.stream),
                          "upload FWI receivers");
                    check(hipMemcpyAsync(w.observed, sh.observed, record * sizeof(float), hipMemcpyHostToDevice,
                                         w.stream),
                          "upload observed data");
                    check(hipStreamSynchronize(w.stream), "upload FWI shot");
                    active.push_back(&w);
                }
                job.run_wave(active, total);
                for (auto* w : active) {
                    unsigned int flag = 0;
                    check(hipMemcpy(&flag, w->flags, sizeof(flag), hipMemcpyDeviceToHost), "read FWI flags");
                    if (flag) fail(SEIS3D_ENONFINITE, "Seis3D: non-finite wavefield in the FWI forward pass.");
                    const std::size_t record = std::size_t(f.samples) * w->nrec * 3;
                    host.resize(record);
                    check(hipMemcpy(host.data(), w->residual, record * sizeof(float), hipMemcpyDeviceToHost),
                          "read residual");
                    for (float r : host) sum += 0.5 * double(r) * double(r);
                }
            }
            // Per-cell 1/rho and shear terms into two temporaries, then fold + chain rule.
            ShotWork& w0 = work[0];
            seis3d::AdjStep A = job.adjoint_operands(w0, w0.slots[0], 0);
            A.g_lam = total[0]; A.g_dev = total[1];
            for (int k = 0; k < 3; ++k) { A.g_shear[k] = total[2 + k]; A.g_buoy[k] = total[5 + k]; }
            check(seis3d::launch_adjoint_cells(A, w0.t[0], w0.t[1], w0.stream), "kernel launch");
            seis3d::FoldParams F{};
            F.g = base.g;
            F.nx = c.nx; F.ny = c.ny; F.nz = c.nz; F.ox = c.ox; F.oy = c.oy; F.oz = c.oz;
            F.inv_rho = base.p.inv_rho; F.lambda = base.p.lambda; F.mu = base.p.mu;
            F.g_ib = w0.t[0]; F.g_lam = total[0]; F.g_dev = w0.t[1];
            const std::size_t phys = std::size_t(c.nx) * c.ny * c.nz;
            F.g_vp = w0.t[2]; F.g_vs = w0.t[3]; F.g_rho = w0.t[4];
            if (phys > L.cells) invalid("Seis3D: internal gradient layout error.");
            check(seis3d::launch_adjoint_fold(F, w0.stream), "kernel launch");
            check(hipMemcpyAsync(g_vp, F.g_vp, phys * sizeof(float), hipMemcpyDeviceToHost, w0.stream), "read gradient");
            check(hipMemcpyAsync(g_vs, F.g_vs, phys * sizeof(float), hipMemcpyDeviceToHost, w0.stream), "read gradient");
            check(hipMemcpyAsync(g_rho, F.g_rho, phys * sizeof(float), hipMemcpyDeviceToHost, w0.stream),
                  "read gradient");
            check(hipStreamSynchronize(w0.stream), "read gradient");
        } catch (...) {
            for (auto& w : work) w.release();
            (void)hipFree(totals_mem);
            throw;
        }
        for (auto& w : work) w.release();
        (void)hipFree(totals_mem);
        // The session's own wavefield was not touched; its stream stays usable.
        if (misfit) *misfit = sum;
    });
}

}  // extern "C"
