// A caller written against the reference-named headers only (include/Seis3D.hpp and
// include/Seis2D.hpp), the way an application that used cuda/Cuda_Seis3D.cuh and
// CUDA_Kernels/Cuda_Check.cuh would after switching libraries.  It is a smoke program for the
// drop-in, with checks of its own design:
//
//  host (always):
//    - the 300x50x300 context asks for dt = 1 ms against vp = 4500 m/s and is refused as
//      std::invalid_argument; the automatic step is the whole-microsecond CFL ceiling;
//    - a grid above the node limit is refused.
//  device (exit 77 without a HIP device):
//    - session lifecycle: a second initializeGpu3D on a live handle is std::logic_error,
//      stepping a released handle is std::logic_error, releaseGpu3D may repeat;
//    - a receiver list with one point outside the grid is refused and the old list stays;
//    - mirror symmetry: a centred explosion in a uniform, all-CPML cube of odd size is
//      symmetric under x -> -x, y -> -y and z -> -z, so collocated vx at +d equals -vx at -d
//      (and likewise for y and z), and the three axes see the same trace (the stencil and
//      the CPML do not prefer an axis);
//    - a 2D Type-1 run is deterministic, and its explosive source radiates vz oddly about the
//      source row (single-source mode: a negative active-source count, as
//      runGpuSimulationStep takes it).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "Seis2D.hpp"
#include "Seis3D.hpp"

namespace {

struct Failed : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void expect(bool ok, const std::string& what) {
    if (!ok) throw Failed(what);
}

template <class E, class F>
bool throws(F&& f) {
    try {
        f();
    } catch (const E&) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

void hostChecks() {
    Seis3D::SimulationContext3D c;
    c.nx3D = 300; c.ny3D = 50; c.nz3D = 300; c.pml3D = 20;
    c.maxVp3D = 4500.0f; c.vp3D = 4500.0f; c.vs3D = 2600.0f;
    c.manualTimeStep3D = true; c.requestedDtMs3D = 1.0f;
    expect(throws<std::invalid_argument>([&] { Seis3D::validateContext3D(c); }), "1 ms step above the CFL ceiling accepted");
    c.manualTimeStep3D = false;
    Seis3D::validateContext3D(c);
    const double weights = 1225.0 / 1024 + 245.0 / 3072 + 49.0 / 5120 + 5.0 / 7168;
    const double ceiling = 0.85 * 10.0 / (std::sqrt(3.0) * 4500.0 * weights);
    const double whole_us = std::floor(ceiling * 1e6) * 1e-6;
    expect(std::abs(double(c.dt3D) - whole_us) < 1e-9, "automatic dt is not the floored CFL ceiling");
    std::printf("HOST dt %.0f us (ceiling %.3f us)\n", double(c.dt3D) * 1e6, ceiling * 1e6);

    Seis3D::SimulationContext3D big;
    big.nx3D = 301; big.ny3D = 300; big.nz3D = 400;
    expect(throws<std::invalid_argument>([&] { Seis3D::validateContext3D(big); }), "node limit not enforced");
}

// Collocated receiver traces of a centred explosion in a uniform all-CPML cube.
void mirrorSymmetry3D() {
    Seis3D::SimulationContext3D c;
    c.nx3D = c.ny3D = c.nz3D = 65;
    c.pml3D = 10;
    c.freeSurface3D = false;
    c.modelKind3D = Seis3D::ModelKind3D::Uniform;
    c.vp3D = 3000.0f; c.vs3D = 1700.0f; c.density3D = 2200.0f; c.maxVp3D = 3000.0f;
    c.frequency3D = 25.0f;
    Seis3D::validateContext3D(c);

    Seis3D::GPUSimData3D g;
    Seis3D::initializeGpu3D(g, c);
    expect(throws<std::logic_error>([&] { Seis3D::initializeGpu3D(g, c); }), "second initialization accepted");

    const int m = 32, d = 12;
    const std::vector<Seis3D::Receiver3D> ring = {
        {m + d, m, m}, {m - d, m, m}, {m, m + d, m}, {m, m - d, m}, {m, m, m + d}, {m, m, m - d}};
    Seis3D::setReceivers3D(g, c, ring.data(), int(ring.size()));
    std::vector<Seis3D::Receiver3D> bad = ring;
    bad.push_back({c.nx3D, m, m});
    expect(throws<std::invalid_argument>([&] { Seis3D::setReceivers3D(g, c, bad.data(), int(bad.size())); }),
           "receiver outside the grid accepted");
    expect(g.receiverCount3D == int(ring.size()), "refused receiver list replaced the old one");

    Seis3D::Source3D s;
    s.x3D = s.y3D = s.z3D = m;
    s.kind3D = Seis3D::SourceKind3D::Explosion;
    const int steps = 180;
    std::vector<float> trace(std::size_t(steps) * ring.size() * 3);
    for (int n = 0; n < steps; ++n) {
        Seis3D::runStep3D(g, c, s, std::uint64_t(n));
        const float* v = Seis3D::sampleReceivers3D(g, c);
        std::copy(v, v + ring.size() * 3, trace.begin() + std::size_t(n) * ring.size() * 3);
    }
    Seis3D::releaseGpu3D(g);
    Seis3D::releaseGpu3D(g);
    expect(throws<std::logic_error>([&] { Seis3D::runStep3D(g, c, s, 0); }), "step on a released session accepted");

    auto at = [&](int n, int r, int comp) { return double(trace[(std::size_t(n) * ring.size() + r) * 3 + comp]); };
    double peak = 0, asym = 0, aniso = 0;
    for (int n = 0; n < steps; ++n) {
        for (int axis = 0; axis < 3; ++axis) {
            const double plus = at(n, 2 * axis, axis), minus = at(n, 2 * axis + 1, axis);
            peak = std::max(peak, std::abs(plus));
            asym = std::max(asym, std::abs(plus + minus));
            aniso = std::max(aniso, std::abs(plus - at(n, 0, 0)));
        }
    }
    std::printf("MIRROR peak %.6e antisymmetry %.3e axis spread %.3e\n", peak, asym / peak, aniso / peak);
    expect(peak > 0 && std::isfinite(peak), "no signal reached the receivers");
    expect(asym <= 1e-5 * peak, "radial velocity is not odd under the mirror");
    expect(aniso <= 1e-4 * peak, "the three axes do not see the same radial trace");
}

// 2D Type-1 split PML, explosive source at the centre of a uniform square.
std::vector<float> run2D(int size, int npml, const std::vector<int>& receivers, int nt) {
    SimulationContext c;
    c.NX = c.NZ = size;
    c.total_grid = int(checkedGridElementCount(size, size));
    c.h = 5.0f; c.dt = 5e-4f; c.nt = nt; c.npml = npml; c.flag_type = 1;
    const double weights[4] = {1225.0 / 1024, -245.0 / 3072, 49.0 / 5120, -5.0 / 7168};
    c.c1_h = float(weights[0] / c.h); c.c2_h = float(weights[1] / c.h);
    c.c3_h = float(weights[2] / c.h); c.c4_h = float(weights[3] / c.h);
    const float vp = 3000.0f, vs = 1732.0f, rho = 2000.0f;
    c.rho.assign(std::size_t(c.total_grid), rho);
    c.mu.assign(std::size_t(c.total_grid), rho * vs * vs);
    c.lambda2mu.assign(std::size_t(c.total_grid), rho * vp * vp);
    c.lambda.assign(std::size_t(c.total_grid), rho * vp * vp - 2 * rho * vs * vs);
    updatePmlDampingArrays(c);
    generateRickerWavelet(c.wavelet, nt, c.dt, 25.0f, 0.04f);
    c.src_idx = (size / 2) * size + size / 2;
    c.src_z_idx = size / 2;
    c.src_type = 2;
    c.num_rcv = int(receivers.size());
    c.rcv_grid_idx = receivers;
    validateSimulationContext(c);

    GPUSimData g;
    initializeGpuSimulation(g, c);
    std::vector<float> vx(receivers.size() * std::size_t(nt)), vz(vx.size());
    for (int it = 0; it < nt; ++it) {
        runGpuSimulationStep(g, it, c, -1);   // a negative count selects the single source
        copyReceiverRecordsFromGpu(g, c.num_rcv, nt, it, vx.data(), vz.data());
    }
    releaseGpuSimulation(g);
    return vz;
}

void explosion2D() {
    // The split PML damps vz with the half-cell profile, which is not mirror symmetric, so the
    // square is large enough that nothing returns from the PML within the 0.25 s record.
    const int size = 241, npml = 20, nt = 500, mid = size / 2, d = 30;
    // vz sits half a cell below its node, (x, z + 1/2), and an explosion's vz is odd about the
    // source row: the image of row mid + d is row mid - d - 1, with the opposite sign.  (Along
    // the source row itself vz is a small difference of the two large split halves, where FP32
    // rounding dominates, so the vertical pair is the sharper check.)
    const std::vector<int> rec = {(mid + d) * size + mid, (mid - d - 1) * size + mid};
    const auto a = run2D(size, npml, rec, nt);
    const auto b = run2D(size, npml, rec, nt);
    expect(a == b, "two identical 2D runs differ");
    double peak = 0, asym = 0;
    for (int it = 0; it < nt; ++it) {
        const double p = a[std::size_t(it)], q = a[std::size_t(nt) + it];
        expect(std::isfinite(p) && std::isfinite(q), "non-finite 2D record");
        peak = std::max(peak, std::abs(p));
        asym = std::max(asym, std::abs(p + q));
    }
    std::printf("2D vz peak %.6e up/down antisymmetry %.3e\n", peak, asym / peak);
    expect(peak > 0, "no 2D signal at the receivers");
    expect(asym <= 2e-5 * peak, "vz is not odd about the source row");
}

}  // namespace

int main() {
    try {
        hostChecks();
        if (seis3d_device_count() < 1) {
            std::puts("SKIP: no HIP device (host checks passed)");
            return 77;
        }
        mirrorSymmetry3D();
        explosion2D();
        std::puts("PASS: drop-in session checks");
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "FAIL: %s\n", e.what());
        return 1;
    }
}
