// Boundary regression of the 3D solver through the reference-named C++ API (include/Seis3D.hpp).
//
// Restates the checks of the reference's tests/Seis3DBoundaryRegression.cpp:
//  * size validation accepts 600x200x200 and 300x300x400 and rejects 400x400x300, INT_MAX^3
//    and a negative dimension (:70-77), and prints the memory budgets (:78-82);
//  * four 96^3 runs of 1600 steps with an explosion (:85-88): late/peak sum of |v|^2 over the
//    physical region below 1e-6 for CPML p12 centre, p40 shallow and the layered p40 model, and
//    above 0.1 for reflecting slip walls (:89);
//  * the centred uniform CPML case keeps its face maxima axially symmetric within 1e-3 for the
//    first 400 steps (:51-55).
// CPU checks always run; without a GPU the program exits 77 (skip), as the reference does.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <vector>

#include "Seis3D.hpp"

namespace {

double measureCase(const char* name, int pml, bool layered, bool reflecting, int sourceDepth) {
    Seis3D::SimulationContext3D c;
    c.nx3D = c.ny3D = c.nz3D = 96;
    c.freeSurface3D = false;
    c.reflecting3D = reflecting;
    c.pml3D = pml;
    c.modelKind3D = layered ? Seis3D::ModelKind3D::TwoLayer : Seis3D::ModelKind3D::Uniform;
    Seis3D::validateContext3D(c);
    Seis3D::GPUSimData3D g;
    Seis3D::initializeGpu3D(g, c);
    struct Release { Seis3D::GPUSimData3D& g; ~Release() { Seis3D::releaseGpu3D(g); } } release{g};
    Seis3D::Source3D source;
    source.x3D = c.nx3D / 2; source.y3D = c.ny3D / 2; source.z3D = sourceDepth;
    source.kind3D = Seis3D::SourceKind3D::Explosion;
    double peak = 0, last = 0;
    const std::size_t plane = std::size_t(c.gx3D) * c.gy3D;
    for (int block = 0; block < 8; ++block) {
        Seis3D::runSteps3D(g, source, std::uint64_t(block) * 200, 200);
        const auto vx = Seis3D::downloadField3D(g, c, Seis3D::Field3D::Vx);
        const auto vy = Seis3D::downloadField3D(g, c, Seis3D::Field3D::Vy);
        const auto vz = Seis3D::downloadField3D(g, c, Seis3D::Field3D::Vz);
        double physical = 0;
        float faceMax[6]{};
        for (int z = 4; z < c.gz3D - 4; ++z) for (int y = 4; y < c.gy3D - 4; ++y) for (int x = 4; x < c.gx3D - 4; ++x) {
            const std::size_t i = Seis3D::gridIndex3D(c, x, y, z);
            const float cx = 0.5f * (vx[i] + vx[i - 1]);
            const float cy = 0.5f * (vy[i] + vy[i - c.gx3D]);
            const float cz = 0.5f * (vz[i] + vz[i - plane]);
            const double e = double(cx) * cx + double(cy) * cy + double(cz) * cz;
            if (!std::isfinite(e)) throw std::runtime_error("Non-finite velocity in boundary case");
            const float m = float(std::sqrt(e));
            const int p[3] = {x - c.ox3D, y - c.oy3D, z - c.oz3D};
            const int n[3] = {c.nx3D, c.ny3D, c.nz3D};
            bool inside = true;
            for (int a = 0; a < 3; ++a) {
                if (p[a] < 0) { faceMax[2 * a] = std::max(faceMax[2 * a], m); inside = false; }
                if (p[a] >= n[a]) { faceMax[2 * a + 1] = std::max(faceMax[2 * a + 1], m); inside = false; }
            }
            if (inside) physical += e;
        }
        peak = std::max(peak, physical);
        last = physical;
        const int step = (block + 1) * 200;
        if (!layered && !reflecting && sourceDepth == 48 && step < 400) {
            for (int axis = 1; axis < 3; ++axis) for (int side = 0; side < 2; ++side)
                if (std::abs(faceMax[2 * axis + side] - faceMax[side]) > 1e-3f * std::max(faceMax[side], 1e-20f))
                    throw std::runtime_error("Centered uniform CPML lost axial symmetry at the top/side faces");
        }
        std::printf("%s,%d,%.6f,%.9g,%.9g,%.9g,%.9g,%.9g,%.9g,%.9g\n", name, step, step * c.dt3D, physical,
                    faceMax[0], faceMax[1], faceMax[2], faceMax[3], faceMax[4], faceMax[5]);
        std::fflush(stdout);
    }
    const double ratio = last / std::max(peak, 1e-30);
    std::printf("SUMMARY %s late/peak velocity-square-sum = %.9g\n", name, ratio);
    return ratio;
}

}  // namespace

int main() {
    try {
        for (const auto size : std::vector<std::array<int, 3>>{{600, 200, 200}, {300, 300, 400}}) {
            Seis3D::SimulationContext3D c;
            c.nx3D = size[0]; c.ny3D = size[1]; c.nz3D = size[2];
            Seis3D::validateContext3D(c);
        }
        for (const auto size : std::vector<std::array<int, 3>>{{400, 400, 300}, {2147483647, 2147483647, 2147483647}, {-1, 96, 96}}) {
            Seis3D::SimulationContext3D c;
            c.nx3D = size[0]; c.ny3D = size[1]; c.nz3D = size[2];
            bool rejected = false;
            try { Seis3D::validateContext3D(c); } catch (const std::invalid_argument&) { rejected = true; }
            if (!rejected) throw std::runtime_error("Invalid grid bypassed the total-node limit");
        }
        for (const auto size : std::vector<std::array<int, 3>>{{96, 96, 96}, {160, 160, 160}, {300, 300, 300}, {300, 300, 400}}) {
            Seis3D::SimulationContext3D c;
            c.nx3D = size[0]; c.ny3D = size[1]; c.nz3D = size[2];
            const auto b = Seis3D::estimateMemory3D(c);
            constexpr double gib = 1073741824.0;
            std::printf("BUDGET %d/%d/%d fields %.6f cpml %.6f total %.6f GiB\n", size[0], size[1], size[2],
                        b.fieldBytes3D / gib, b.cpmlBytes3D / gib, b.totalBytes3D / gib);
        }
        if (seis3d_device_count() < 1) { std::puts("SKIP: no HIP device"); return 77; }
        std::puts("case,step,time_s,physical_v2_sum,x_min,x_max,y_min,y_max,z_min,z_max");
        const double center = measureCase("uniform-center-p12", 12, false, false, 48);
        const double shallow = measureCase("uniform-shallow-p40", 40, false, false, 24);
        const double layers = measureCase("layers-shallow-p40", 40, true, false, 24);
        const double reflecting = measureCase("uniform-reflecting", 12, false, true, 48);
        if (center > 1e-6 || shallow > 1e-6 || layers > 1e-6 || reflecting < .1)
            throw std::runtime_error("Absorbing/reflection late-time contrast regressed");
        std::puts("PASS: total-node validation, finite six-face CPML decay, lossless reflecting reference");
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "FAIL: %s\n", e.what());
        return 1;
    }
}
