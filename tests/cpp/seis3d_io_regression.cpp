// Shot-record and model files through the reference-named C++ API (include/Seis3D.hpp).
//
// Follows the reference's workspace checks (tests/Seis3DWorkspaceRegression.cpp):
//  * a frozen gather of the default survey exported as three SEG-Y rev1 components plus the
//    geometry CSV and read back: samples bit-exact, coordinates to the millimetre scalar, the
//    first-sample time through the time scalar, trace ids, units, rank receivers (:688-702);
//  * a second export to the same prefix is refused and leaves the files unchanged;
//  * a .s3dm model written, inspected (modelKind Imported, max Vp from the payload) and, with a
//    GPU, imported into a session and exported again (:187-192): the same header bytes, and
//    values within a few ulp, since a session stores 1/rho, lambda, mu and the export rebuilds
//    Vp, Vs, rho from them.
// CPU checks always run; without a GPU the program exits 77 (skip), as the reference does.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

#include <unistd.h>

#include "Seis3D.hpp"

namespace fs = std::filesystem;

namespace {

void require(bool ok, const char* what) {
    if (!ok) throw std::runtime_error(what);
}

std::vector<char> bytes(const fs::path& p) {
    std::ifstream in(p, std::ios::binary);
    return std::vector<char>(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
}

void gatherRoundTrip(const fs::path& dir) {
    Seis3D::SimulationContext3D c;
    c.nx3D = 120; c.ny3D = 60; c.nz3D = 80;
    Seis3D::validateContext3D(c);
    Seis3D::Survey3D q;
    q.worldOrigin3D[0] = 612345.25; q.worldOrigin3D[1] = 4123456.5;
    Seis3D::fitSurvey3D(c, q);
    const auto receivers = Seis3D::configureReceivers3D(c, q);
    require(receivers.size() == std::size_t(q.receiverCountX3D) * q.receiverCountY3D, "receiver spread size");
    Seis3D::Source3D s;
    s.x3D = c.nx3D / 2; s.y3D = c.ny3D / 2; s.z3D = c.nz3D / 4;
    auto g = Seis3D::startGather3D(c, s, q, receivers, 7, 64, 3);
    require(g.sampleIntervalUs3D == int(std::llround(double(c.dt3D) * 1e6)) * 3, "sample interval");
    require(std::abs(g.firstSampleTime3D - 2.5 * c.dt3D) < 1e-12, "first sample time (stride - 1/2) dt");
    require(g.coordinateScalar3D == -100, "XY scalar keeps 4.1e6 m northings in range");
    g.sampleCount3D = 64;
    g.samples3D.resize(std::size_t(64) * receivers.size() * 3);
    for (std::size_t i = 0; i < g.samples3D.size(); ++i)
        g.samples3D[i] = float(std::sin(0.37 * double(i)) * 1e-6 * double(i % 97 + 1));
    Seis3D::exportGather3D(g, dir.string(), "line");
    const char* names[3] = {"line_0007_vx.sgy", "line_0007_vy.sgy", "line_0007_vz.sgy"};
    for (int comp = 0; comp < 3; ++comp) {
        const auto shot = Seis3D::readShotSegy3D((dir / names[comp]).string());
        require(shot.component3D == comp && shot.amplitudeUnit3D == "m/s" && shot.cartesian3D, "component and units");
        require(shot.gather3D.shotId3D == 7 && shot.gather3D.sampleIntervalUs3D == g.sampleIntervalUs3D, "shot header");
        const std::size_t n = receivers.size();
        for (std::size_t k = 0; k < 64; ++k)
            for (std::size_t r = 0; r < n; ++r) {
                const std::size_t i = (k * n + r) * 3 + std::size_t(comp);
                require(shot.gather3D.samples3D[i] == g.samples3D[i], "samples bit-exact");
            }
        const double h = c.spacing3D;
        for (std::size_t r = 0; r < n; ++r) {
            require(std::abs(shot.coordinates3D[r][0] - (q.worldOrigin3D[0] + receivers[r].x3D * h)) <= 0.005 + 1e-9,
                    "receiver X");
            require(std::abs(shot.coordinates3D[r][1] - (q.worldOrigin3D[1] + receivers[r].y3D * h)) <= 0.005 + 1e-9,
                    "receiver Y");
            require(shot.traceIds3D[r] == int(r) + 1, "trace ids");
            require(std::abs(shot.firstTimes3D[r] - g.firstSampleTime3D) <= 0.5e-7 + 1e-12, "time scalar");
        }
        // Receivers are coordinate ranks: the default spread is a regular grid, so ranks count it.
        require(shot.gather3D.receivers3D[n - 1].x3D == q.receiverCountX3D - 1 &&
                    shot.gather3D.receivers3D[n - 1].y3D == q.receiverCountY3D - 1,
                "rank receivers");
    }
    require(fs::exists(dir / "line_0007_geometry.csv"), "geometry manifest");
    const auto before = bytes(dir / names[0]);
    bool refused = false;
    try {
        Seis3D::exportGather3D(g, dir.string(), "line");
    } catch (const std::invalid_argument&) {
        refused = true;
    }
    require(refused && bytes(dir / names[0]) == before, "second export refused, files unchanged");
    std::printf("GATHER %zu receivers x 64 samples, interval %d us, scalar %d\n", receivers.size(),
                g.sampleIntervalUs3D, g.coordinateScalar3D);
}

Seis3D::SimulationContext3D modelFile(const fs::path& path) {
    Seis3D::SimulationContext3D c;
    c.nx3D = 40; c.ny3D = 32; c.nz3D = 24; c.pml3D = 10;
    Seis3D::validateContext3D(c);
    std::vector<float> planes(std::size_t(c.nx3D) * c.ny3D * c.nz3D * 3);
    for (int z = 0; z < c.nz3D; ++z)
        for (int y = 0; y < c.ny3D; ++y)
            for (int x = 0; x < c.nx3D; ++x) {
                const std::size_t i = ((std::size_t(z) * c.ny3D + y) * c.nx3D + x) * 3;
                planes[i] = 2000.0f + 50.0f * float(z) + float((x * 7 + y * 3) % 11);
                planes[i + 1] = planes[i] / 1.8f;
                planes[i + 2] = 2100.0f + float(z);
            }
    const seis3d_ctx cc = Seis3D::detail::toC(c);
    Seis3D::detail::check(seis3d_s3dm_write(path.string().c_str(), &cc, planes.data()));
    const auto m = Seis3D::inspectModel3D(path.string(), c);
    require(m.modelKind3D == Seis3D::ModelKind3D::Imported && m.nx3D == 40 && m.ny3D == 32 && m.nz3D == 24, "inspect");
    require(m.maxVp3D == 2000.0f + 50.0f * 23 + 10.0f, "max Vp from the payload");
    std::printf("MODEL %dx%dx%d max Vp %.1f dt %.6f\n", m.nx3D, m.ny3D, m.nz3D, m.maxVp3D, m.dt3D);
    return m;
}

}  // namespace

int main() {
    const fs::path dir = fs::temp_directory_path() / ("seis3d_io_regression_" + std::to_string(::getpid()));
    fs::remove_all(dir);
    fs::create_directories(dir);
    int rc = 0;
    try {
        gatherRoundTrip(dir / "shots");
        const auto m = modelFile(dir / "model.s3dm");
        if (seis3d_device_count() < 1) {
            std::puts("SKIP: no HIP device");
            rc = 77;
        } else {
            Seis3D::GPUSimData3D gpu;
            Seis3D::initializeGpu3D(gpu, m);
            Seis3D::importModel3D(gpu, m, (dir / "model.s3dm").string());
            Seis3D::exportModel3D(gpu, m, (dir / "copy.s3dm").string());
            Seis3D::releaseGpu3D(gpu);
            const auto a = bytes(dir / "model.s3dm"), b = bytes(dir / "copy.s3dm");
            require(a.size() == b.size() && std::equal(a.begin(), a.begin() + 36, b.begin()), "model header");
            for (std::size_t i = 36; i + 4 <= a.size(); i += 4) {
                float x, y;
                std::memcpy(&x, a.data() + i, 4);
                std::memcpy(&y, b.data() + i, 4);
                require(std::abs(x - y) <= 1e-6f * std::abs(x), "model values within a few ulp");
            }
            std::puts("PASS: SEG-Y rev1 gather round trip, export collision refusal, .s3dm GPU round trip");
        }
    } catch (const std::exception& e) {
        std::printf("FAIL: %s\n", e.what());
        rc = 1;
    }
    fs::remove_all(dir);
    return rc;
}
