// 2D shot records and packages through the reference-named C++ API (include/Seis2D.hpp).
//
// Restates the export checks of the reference's tests/Seis2DRecordRegression.cpp:69-97:
//  * a two-trace record exported and read back unchanged, in m/s; an invalid interval refused
//    without touching the existing file;
//  * a package with the frozen 8x8 model in a Unicode directory: raw data unchanged, model.s2dm
//    of 20 + 64 * 12 bytes, the model magic escaped in manifest.json, a second export refused
//    without overwriting, unsafe prefixes refused, a NaN shot leaving no folder, no staging
//    folders left behind.
// Host-only: runs (and must pass) everywhere.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "Seis2D.hpp"

namespace fs = std::filesystem;

namespace {

void require(bool ok, const char* what) {
    if (!ok) throw std::runtime_error(what);
}

}  // namespace

int main() {
    const fs::path dir = fs::temp_directory_path() /
                         ("seis2d_record_" + std::to_string(std::chrono::steady_clock::now().time_since_epoch().count()));
    int rc = 0;
    try {
        fs::create_directories(dir);
        std::vector<SeismicIO::TraceMetadata> geometry(2);
        const std::vector<std::vector<float>> values = {{0, 1, -2}, {3, 0, -1}};
        const fs::path exported = dir / "export.sgy";
        SeismicIO::writeSegyCaptureFile2D(values, geometry, exported, .001f);
        const auto roundtrip = Seis2D::readRecordSegy2D(exported);
        require(roundtrip.traces == values && roundtrip.units == "m/s", "export round trip changed data or units");
        bool rejected = false;
        try {
            SeismicIO::writeSegyCaptureFile2D(values, geometry, exported, .1f);
        } catch (const std::invalid_argument&) {
            rejected = true;
        }
        require(rejected && Seis2D::readRecordSegy2D(exported).traces == values, "invalid export overwrote a record");

        Seis2D::FrozenModel2D frozen;
        frozen.nx = frozen.nz = 8; frozen.spacing = 10; frozen.solverDt = .001f;
        frozen.formulation = 1; frozen.frequency = 20; frozen.sourceAmplitude = 1;
        frozen.lambda.assign(64, 2e9f); frozen.mu.assign(64, 1e9f); frozen.rho.assign(64, 2000); frozen.wavelet = {0, 1, 0};
        Seis2D::ShotExportRequest2D package;
        package.directory = dir / fs::u8path(u8"中文归档");
        package.prefix = "shot"; package.samples = 3; package.sampleInterval = .001f; package.version = "regression";
        geometry[0].coordinateScalar = geometry[1].coordinateScalar = -1000;
        geometry[0].sourceX = 12345;
        geometry[0].traceInRecord = 1; geometry[1].traceInRecord = 2;
        const auto published = Seis2D::exportShotPackage2D(package, frozen, values, values, geometry);
        require(Seis2D::readRecordSegy2D(published / "record_vz.sgy").traces == values, "package changed raw data");
        require(fs::file_size(published / "model.s2dm") == 20 + 64 * 12 && fs::exists(published / "manifest.json"),
                "frozen model or manifest missing");
        std::ifstream manifestFile(published / "manifest.json");
        const std::string manifest((std::istreambuf_iterator<char>(manifestFile)), std::istreambuf_iterator<char>());
        require(manifest.find("S2DMOD1\\\\0") != std::string::npos, "model magic not escaped as JSON text");
        bool collision = false;
        try {
            Seis2D::exportShotPackage2D(package, frozen, values, values, geometry);
        } catch (const std::runtime_error&) {
            collision = true;
        }
        require(collision && Seis2D::readRecordSegy2D(published / "record_vx.sgy").traces == values,
                "package collision overwrote raw data");
        for (const char* prefix : {"../escape", "CON", "bad:name", "trailing."}) {
            bool bad = false;
            try {
                Seis2D::validateExportPrefix2D(prefix);
            } catch (const std::invalid_argument&) {
                bad = true;
            }
            require(bad, "unsafe package prefix accepted");
        }
        package.shot = 2;
        auto invalid = values;
        invalid[0][0] = std::numeric_limits<float>::quiet_NaN();
        bool refused = false;
        try {
            Seis2D::exportShotPackage2D(package, frozen, invalid, values, geometry);
        } catch (const std::invalid_argument&) {
            refused = true;
        }
        require(refused && !fs::exists(package.directory / "shot_0002"), "invalid package created output");
        for (const auto& item : fs::directory_iterator(package.directory))
            require(item.path().filename().string().find(".part-") == std::string::npos, "staging folder leaked");
        std::puts("PASS: record round trip, Unicode shot package, frozen model, manifest/SHA-256, collision and "
                  "invalid-input isolation");
    } catch (const std::exception& e) {
        std::printf("FAIL: %s\n", e.what());
        rc = 1;
    }
    std::error_code ignored;
    fs::remove_all(dir, ignored);
    return rc;
}
