// seis3d_io.cpp - shot-record and model files of the 3D solver (host only).
//
//  * Gather metadata as startGather3D (src/Seis3D/Acquisition3D.cpp:188-230).
//  * SEG-Y rev1 export of the three velocity components plus a geometry CSV
//    (writeSegyComponent3D / exportGather3D, :40-185): 3200-byte EBCDIC-free
//    ASCII text header, 400-byte binary header, 240-byte trace headers and
//    big-endian IEEE float samples, with the reference's header byte map
//    (docs/development/SEIS3D_IMPLEMENTATION.md:212-225) so files interoperate.
//  * SEG-Y rev0/rev1 import of one single-component shot, sample formats
//    1 (IBM), 2, 3, 5 (IEEE) and 8 (readShotSegy3D, ShotRecordIo3D.cpp:66-141).
//  * The .s3dm model format (ModelIo3D.cpp:35-102): 36-byte little-endian
//    header, then float32 [Vp, Vs, rho] planes, x fastest.
// Files are written through "<name>.part" and renamed; existing files are
// never overwritten.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/seis3d.h"
#include "seis3d_host.hpp"

using namespace seis3d::host;
namespace fs = std::filesystem;

namespace {

void put16(unsigned char* o, int at, std::uint16_t v) {
    o[at] = static_cast<unsigned char>(v >> 8);
    o[at + 1] = static_cast<unsigned char>(v);
}
void put32(unsigned char* o, int at, std::int32_t v) {
    const auto u = static_cast<std::uint32_t>(v);
    for (int k = 0; k < 4; ++k) o[at + k] = static_cast<unsigned char>(u >> (24 - 8 * k));
}
unsigned get16(const unsigned char* p) { return (unsigned(p[0]) << 8) | p[1]; }
int get16s(const unsigned char* p) {
    const unsigned v = get16(p);
    return v < 32768 ? int(v) : int(v) - 65536;
}
std::uint32_t get32u(const unsigned char* p) {
    return (std::uint32_t(p[0]) << 24) | (std::uint32_t(p[1]) << 16) | (std::uint32_t(p[2]) << 8) | p[3];
}
std::int32_t get32(const unsigned char* p) {
    const std::uint32_t u = get32u(p);
    std::int32_t v;
    std::memcpy(&v, &u, 4);
    return v;
}
// SEG-Y signed scalar: negative divides, positive multiplies, zero means one.
double scale_of(int s) { return s < 0 ? 1.0 / -double(s) : s > 0 ? double(s) : 1.0; }

std::int32_t millimetres(double metres) {
    const double mm = std::round(metres * 1000);
    if (!std::isfinite(mm) || mm < std::numeric_limits<std::int32_t>::min() || mm > std::numeric_limits<std::int32_t>::max())
        fail(SEIS3D_EINVAL, "SEG-Y coordinate exceeds the signed 32-bit millimetre range.");
    return std::int32_t(mm);
}
std::int32_t scaled(double metres, int scalar) {
    const double v = std::round(metres * (scalar < 0 ? -scalar : 1.0 / scalar));
    if (!std::isfinite(v) || v < std::numeric_limits<std::int32_t>::min() || v > std::numeric_limits<std::int32_t>::max())
        fail(SEIS3D_EINVAL, "SEG-Y world coordinate exceeds the supported integer range.");
    return std::int32_t(v);
}

// One-based CMP (inline, crossline) bins of a source-receiver midpoint: inline bins Y, crossline bins X.
std::pair<int, int> cmp_bin(const seis3d_gather_info& g, const seis3d_receiver& r) {
    const double x = 0.5 * (g.source.x + r.x) * g.spacing, y = 0.5 * (g.source.y + r.y) * g.spacing;
    const double il = std::floor(y / g.cmp_bin[1]) + 1, xl = std::floor(x / g.cmp_bin[0]) + 1;
    if (!std::isfinite(il) || !std::isfinite(xl) || il > 2147483646.0 || xl > 2147483646.0)
        fail(SEIS3D_EINVAL, "CMP bin index overflow.");
    return {int(il), int(xl)};
}

std::string fixed(double v, int digits) {
    std::ostringstream s;
    s << std::fixed << std::setprecision(digits) << v;
    return s.str();
}

void write_component(const seis3d_gather_info& g, const seis3d_receiver* rec, int count, const float* samples,
                     const fs::path& path, int comp) {
    const fs::path part = fs::path(path.string() + ".part");
    try {
        std::ofstream out(part, std::ios::binary | std::ios::trunc);
        if (!out) fail(SEIS3D_EHIP, "Cannot create SEG-Y file " + part.string());
        std::array<char, 3200> text;
        text.fill(' ');
        const char* names[3] = {"VX", "VY", "VZ"};
        const std::string lines[] = {
            "3D ISOTROPIC ELASTIC SHOT GATHER, SEG-Y REV1, ASCII TEXT HEADER",
            std::string("COMPONENT ") + names[comp] + ", PARTICLE VELOCITY M/S, IEEE FLOAT32 BIG-ENDIAN",
            "COORDINATES IN METRES, XY SCALAR " + std::to_string(g.coordinate_scalar) + ", ELEVATION SCALAR -1000",
            "TRACE BYTES 181-188 MIDPOINT XY, 189-192 CMP INLINE, 193-196 CMP CROSSLINE",
            "CMP INLINE BINS Y, CROSSLINE BINS X, ONE-BASED; Z POSITIVE DOWN",
            "FIRST SAMPLE TIME S " + fixed(g.first_sample_time, 9) + " (TIME SCALAR AT BYTES 215-216)",
            "VELOCITIES AT LEAPFROG HALF STEPS; THE GEOMETRY CSV HOLDS EXACT TIMES",
            "SHOTPOINT BYTES 197-200 = SOURCE ID",
            "GRID SPACING M " + fixed(g.spacing, 3) + ", SOURCE PEAK FREQUENCY HZ " + fixed(g.frequency, 3),
            "SOURCE AMPLITUDE " + fixed(g.source.amplitude, 3) +
                (g.source.kind == SEIS3D_EXPLOSION ? " PA/S EXPLOSION" : " N/M3 BODY FORCE")};
        for (int line = 0; line < 40; ++line) {
            char label[5]{};
            std::snprintf(label, sizeof(label), "C%02d ", line + 1);
            std::memcpy(text.data() + line * 80, label, 4);
        }
        for (std::size_t i = 0; i < std::size(lines); ++i)
            std::memcpy(text.data() + i * 80 + 4, lines[i].data(), std::min<std::size_t>(76, lines[i].size()));
        out.write(text.data(), text.size());
        std::array<unsigned char, 400> bin{};
        put32(bin.data(), 0, 1);
        put32(bin.data(), 4, 1);
        put16(bin.data(), 12, std::uint16_t(count));
        put16(bin.data(), 16, std::uint16_t(g.sample_interval_us));
        put16(bin.data(), 18, std::uint16_t(g.sample_interval_us));
        put16(bin.data(), 20, std::uint16_t(g.samples));
        put16(bin.data(), 22, std::uint16_t(g.samples));
        put16(bin.data(), 24, 5);        // IEEE float
        put16(bin.data(), 28, 5);        // common-source ensemble sorting
        put16(bin.data(), 54, 1);        // metres
        put16(bin.data(), 300, 0x0100);  // rev1
        put16(bin.data(), 302, 1);       // fixed-length traces
        out.write(reinterpret_cast<const char*>(bin.data()), bin.size());
        const double h = g.spacing;
        const double sx = g.world_origin[0] + g.source.x * h, sy = g.world_origin[1] + g.source.y * h;
        int tf = 10000;
        while (tf > 1 && std::round(g.first_sample_time * 1000 * tf) > 32767) tf /= 10;
        std::vector<unsigned char> data(std::size_t(g.samples) * 4);
        for (int r = 0; r < count; ++r) {
            const seis3d_receiver p = rec[r];
            const double gx = g.world_origin[0] + p.x * h, gy = g.world_origin[1] + p.y * h;
            const auto bin2 = cmp_bin(g, p);
            std::array<unsigned char, 240> t{};
            put32(t.data(), 0, r + 1);
            put32(t.data(), 4, r + 1);
            put32(t.data(), 8, g.shot_id);
            put32(t.data(), 12, r + 1);
            put32(t.data(), 16, g.shot_id);
            put32(t.data(), 20, g.shot_id);
            put32(t.data(), 24, r + 1);
            put16(t.data(), 28, std::uint16_t(comp == 0 ? 14 : comp == 1 ? 13 : 12));
            put16(t.data(), 30, 1);
            put16(t.data(), 32, 1);
            put32(t.data(), 36, std::int32_t(std::round(std::hypot(gx - sx, gy - sy))));
            put32(t.data(), 40, millimetres(-p.z * h));
            put32(t.data(), 44, 0);
            put32(t.data(), 48, millimetres(g.source.z * h));
            put16(t.data(), 68, std::uint16_t(std::int16_t(-1000)));
            put16(t.data(), 70, std::uint16_t(std::int16_t(g.coordinate_scalar)));
            put32(t.data(), 72, scaled(sx, g.coordinate_scalar));
            put32(t.data(), 76, scaled(sy, g.coordinate_scalar));
            put32(t.data(), 80, scaled(gx, g.coordinate_scalar));
            put32(t.data(), 84, scaled(gy, g.coordinate_scalar));
            put16(t.data(), 88, 1);
            put16(t.data(), 108, std::uint16_t(std::round(g.first_sample_time * 1000 * tf)));
            put16(t.data(), 114, std::uint16_t(g.samples));
            put16(t.data(), 116, std::uint16_t(g.sample_interval_us));
            put32(t.data(), 180, scaled(0.5 * (sx + gx), g.coordinate_scalar));
            put32(t.data(), 184, scaled(0.5 * (sy + gy), g.coordinate_scalar));
            put32(t.data(), 188, bin2.first);
            put32(t.data(), 192, bin2.second);
            put32(t.data(), 196, g.shot_id);
            put16(t.data(), 200, 1);
            put16(t.data(), 202, 6);   // m/s
            put32(t.data(), 204, 1);
            put16(t.data(), 210, 6);
            put16(t.data(), 212, std::uint16_t(r + 1));
            put16(t.data(), 214, std::uint16_t(std::int16_t(-tf)));
            for (int k = 0; k < g.samples; ++k) {
                const float v = samples[(std::size_t(k) * count + r) * 3 + comp];
                if (!std::isfinite(v)) fail(SEIS3D_ENONFINITE, "Refusing to export non-finite receiver samples.");
                std::uint32_t bits;
                std::memcpy(&bits, &v, 4);
                put32(data.data(), k * 4, std::int32_t(bits));
            }
            out.write(reinterpret_cast<const char*>(t.data()), t.size());
            out.write(reinterpret_cast<const char*>(data.data()), std::streamsize(data.size()));
            if (!out) fail(SEIS3D_EHIP, "SEG-Y write failed; check disk space.");
        }
        out.close();
        if (!out) fail(SEIS3D_EHIP, "Cannot finalize SEG-Y file.");
        if (fs::exists(path)) fail(SEIS3D_EINVAL, "SEG-Y destination already exists.");
        fs::rename(part, path);
    } catch (...) {
        std::error_code ignored;
        fs::remove(part, ignored);
        throw;
    }
}

std::string shot_base(const char* prefix, int shot) {
    const std::string p = prefix ? prefix : "";
    if (p.empty() || p.size() > 120 || p.find_first_of("/\\:*?\"<>|") != std::string::npos || p == "." || p == "..")
        invalid("Seis3D: invalid export prefix.");
    std::ostringstream s;
    s << p << '_' << std::setw(4) << std::setfill('0') << shot;
    return s.str();
}

// ---- SEG-Y import ----

struct Reader {
    std::ifstream in;
    std::uint64_t size = 0, begin = 0, record = 0;
    std::array<unsigned char, 400> bin{};
    std::array<unsigned char, 240> first{};
    seis3d_segy_info info{};
    int bytes = 4;

    void exact(void* out, std::size_t n) {
        if (!in.read(static_cast<char*>(out), std::streamsize(n))) fail(SEIS3D_EINVAL, "Truncated SEG-Y header or trace samples.");
    }

    explicit Reader(const char* path) {
        if (!path) invalid("Seis3D: null path.");
        std::error_code ec;
        size = fs::file_size(path, ec);
        if (ec) fail(SEIS3D_EINVAL, "Cannot open shot SEG-Y file.");
        in.open(path, std::ios::binary);
        if (!in) fail(SEIS3D_EINVAL, "Cannot open shot SEG-Y file.");
        if (size < 3840) fail(SEIS3D_EINVAL, "SEG-Y is too small to contain a trace.");
        in.seekg(3200);
        exact(bin.data(), bin.size());
        info.revision = int(get16(bin.data() + 300));
        if (info.revision != 0 && info.revision != 256) fail(SEIS3D_EINVAL, "Supported SEG-Y revisions: big-endian rev0 and rev1.");
        info.format = int(get16(bin.data() + 24));
        const int f = info.format;
        if (f != 1 && f != 2 && f != 3 && f != 5 && f != 8) fail(SEIS3D_EINVAL, "Unsupported SEG-Y sample format.");
        bytes = f == 3 ? 2 : f == 8 ? 1 : 4;
        const int extended = info.revision ? get16s(bin.data() + 304) : 0;
        if (extended < 0) fail(SEIS3D_EINVAL, "Variable-count extended text headers are not supported.");
        begin = 3600 + std::uint64_t(extended) * 3200;
        if (begin + 240 > size) fail(SEIS3D_EINVAL, "SEG-Y extended headers exceed file length.");
        in.seekg(std::streamoff(begin));
        exact(first.data(), first.size());
        info.samples = get16(first.data() + 114) ? int(get16(first.data() + 114)) : int(get16(bin.data() + 20));
        info.sample_interval_us = get16(first.data() + 116) ? int(get16(first.data() + 116)) : int(get16(bin.data() + 16));
        if (info.samples < 1 || info.samples > kMaxSegySamples || info.sample_interval_us < 1)
            fail(SEIS3D_EINVAL, "SEG-Y requires 1..32767 samples per trace and a positive sample interval.");
        record = 240 + std::uint64_t(info.samples) * bytes;
        const std::uint64_t payload = size - begin;
        if (payload % record) fail(SEIS3D_EINVAL, "SEG-Y trace lengths are inconsistent or the file is truncated.");
        const std::uint64_t traces = payload / record;
        if (traces < 1 || traces > 4096 || traces * std::uint64_t(info.samples) * 3 * sizeof(float) > 256ull * 1024 * 1024)
            fail(SEIS3D_EINVAL, "Imported shot exceeds 4096 traces or the 256 MiB sample budget.");
        info.traces = int(traces);
        info.shot_id = get32(first.data() + 8);
        const int id = int(get16(first.data() + 28));
        info.component = id == 14 ? 0 : id == 13 ? 1 : 2;
        info.measurement_unit = info.revision ? get16s(first.data() + 202) : 0;
        info.coordinate_units = int(get16(first.data() + 88));
    }

    float sample(const unsigned char* p) const {
        double v = 0;
        const int f = info.format;
        if (f == 5) {
            const std::uint32_t b = get32u(p);
            float x;
            std::memcpy(&x, &b, 4);
            v = x;
        } else if (f == 1) {
            const std::uint32_t b = get32u(p);
            v = std::ldexp(double(b & 0xffffff) / 16777216.0, 4 * (int((b >> 24) & 127) - 64));
            if (b >> 31) v = -v;
        } else if (f == 2) {
            v = get32(p);
        } else if (f == 3) {
            v = get16s(p);
        } else {
            v = p[0] < 128 ? int(p[0]) : int(p[0]) - 256;
        }
        if (!std::isfinite(v) || std::abs(v) > std::numeric_limits<float>::max())
            fail(SEIS3D_EINVAL, "SEG-Y contains non-finite or overflowing samples.");
        return float(v);
    }

    void read_all(float* samples, double* coords, double* times, std::int32_t* ids) {
        const int n = info.traces, ns = info.samples;
        const int length_units = int(get16(bin.data() + 54));
        const double length_factor = length_units == 2 ? 0.3048 : 1.0;
        std::vector<unsigned char> data(std::size_t(ns) * bytes);
        std::array<double, 2> src{};
        in.seekg(std::streamoff(begin));
        for (int r = 0; r < n; ++r) {
            std::array<unsigned char, 240> t{};
            exact(t.data(), t.size());
            const int tns = get16(t.data() + 114) ? int(get16(t.data() + 114)) : ns;
            const int tdt = get16(t.data() + 116) ? int(get16(t.data() + 116)) : info.sample_interval_us;
            if (tns != ns || tdt != info.sample_interval_us)
                fail(SEIS3D_EINVAL, "Variable trace lengths or sample intervals are not supported.");
            if (get32(t.data() + 8) != info.shot_id) fail(SEIS3D_EINVAL, "File contains multiple shots. Import one shot at a time.");
            if (get32(t.data() + 16) != get32(first.data() + 16))
                fail(SEIS3D_EINVAL, "Energy source identifiers differ; expected one shot.");
            if (get16(t.data() + 28) != get16(first.data() + 28) ||
                (info.revision ? get16s(t.data() + 202) : 0) != info.measurement_unit)
                fail(SEIS3D_EINVAL, "Mixed trace components/units are not supported in a single imported file.");
            if (int(get16(t.data() + 88)) != info.coordinate_units) fail(SEIS3D_EINVAL, "Mixed coordinate units in SEG-Y traces.");
            const double cs = scale_of(get16s(t.data() + 70)) * (info.coordinate_units == 1 ? length_factor : 1.0);
            const double es = scale_of(get16s(t.data() + 68)) * length_factor;
            const std::array<double, 2> here{get32(t.data() + 72) * cs, get32(t.data() + 76) * cs};
            if (r && here != src) fail(SEIS3D_EINVAL, "Source coordinates differ between traces; expected one common shot.");
            src = here;
            if (coords) {
                coords[3 * r] = get32(t.data() + 80) * cs;
                coords[3 * r + 1] = get32(t.data() + 84) * cs;
                coords[3 * r + 2] = -get32(t.data() + 40) * es;   // depth below the datum
            }
            if (times) times[r] = get16s(t.data() + 108) * 0.001 * (info.revision ? scale_of(get16s(t.data() + 214)) : 1.0);
            if (ids) {
                const int id = get32(t.data() + 12);
                ids[r] = id > 0 ? id : r + 1;
            }
            exact(data.data(), data.size());
            for (int k = 0; k < ns; ++k) {
                const float v = sample(data.data() + std::size_t(k) * bytes);
                if (samples) samples[std::size_t(r) * ns + k] = v;
            }
        }
    }
};

// ---- .s3dm ----

std::uint32_t le32(std::istream& in) {
    unsigned char b[4]{};
    in.read(reinterpret_cast<char*>(b), 4);
    if (!in) fail(SEIS3D_EINVAL, "Seis3D model file is truncated.");
    return std::uint32_t(b[0]) | (std::uint32_t(b[1]) << 8) | (std::uint32_t(b[2]) << 16) | (std::uint32_t(b[3]) << 24);
}
void put_le32(std::ostream& out, std::uint32_t v) {
    const unsigned char b[4] = {static_cast<unsigned char>(v), static_cast<unsigned char>(v >> 8),
                                static_cast<unsigned char>(v >> 16), static_cast<unsigned char>(v >> 24)};
    out.write(reinterpret_cast<const char*>(b), 4);
}

template <class PlaneSource>
void write_model(const char* path, const seis3d_ctx& c, PlaneSource&& plane_of) {
    if (!path || !*path) invalid("Seis3D: choose a model filename.");
    const fs::path dst(path);
    if (fs::exists(dst)) invalid("Choose a new model filename; the destination already exists.");
    if (!dst.parent_path().empty()) fs::create_directories(dst.parent_path());
    const fs::path part = fs::path(dst.string() + ".part");
    try {
        std::ofstream out(part, std::ios::binary | std::ios::trunc);
        if (!out) fail(SEIS3D_EHIP, "Cannot create Seis3D model file.");
        out.write("S3DMOD1", 8);
        put_le32(out, c.reflecting ? 2 : 1);
        put_le32(out, std::uint32_t(c.nx));
        put_le32(out, std::uint32_t(c.ny));
        put_le32(out, std::uint32_t(c.nz));
        put_le32(out, std::uint32_t(c.pml));
        put_le32(out, c.reflecting ? 2 : c.free_surface ? 1 : 0);
        std::uint32_t spacing = 0;
        std::memcpy(&spacing, &c.spacing, 4);
        put_le32(out, spacing);
        std::vector<float> plane(std::size_t(c.nx) * c.ny * 3);
        for (int z = 0; z < c.nz; ++z) {
            plane_of(z, plane.data());
            out.write(reinterpret_cast<const char*>(plane.data()), std::streamsize(plane.size() * 4));
            if (!out) fail(SEIS3D_EHIP, "Seis3D model write failed; check disk space.");
        }
        out.close();
        if (!out) fail(SEIS3D_EHIP, "Cannot finalize model file.");
        fs::rename(part, dst);
    } catch (...) {
        std::error_code ignored;
        fs::remove(part, ignored);
        throw;
    }
}

}  // namespace

extern "C" {

int seis3d_gather_info_init(const seis3d_ctx* ctx, const seis3d_source* src, const seis3d_receiver* rec, int count,
                            int shot_id, int stride, int samples, const double world_origin[2], const float cmp_bin_size[2],
                            seis3d_gather_info* out) {
    return guarded([&] {
        if (!ctx || !src || !out || !world_origin || !cmp_bin_size || (count > 0 && !rec)) invalid("Seis3D: invalid gather request.");
        const seis3d_ctx& c = *ctx;
        if (src->x < 1 || src->x >= c.nx - 1 || src->y < 1 || src->y >= c.ny - 1 || src->z < 2 || src->z >= c.nz - 1 ||
            !std::isfinite(src->amplitude))
            invalid("Place the source inside the model, at depth >=2, before capture.");
        for (float d : src->direction)
            if (!std::isfinite(d)) invalid("Invalid source direction.");
        if (count < 1 || count > kMaxReceivers) invalid("Seis3D: a gather needs 1..300 receivers (maxReceivers3D).");
        if (samples < 1 || samples > kMaxSegySamples || stride < 1 || stride > 10000)
            invalid("Gather sample count must be 1..32767 with stride 1..10000.");
        for (int a = 0; a < 2; ++a)
            if (!std::isfinite(cmp_bin_size[a]) || cmp_bin_size[a] <= 0 || !std::isfinite(world_origin[a]))
                invalid("Invalid CMP bin size or world origin.");
        const long long interval = std::llround(double(c.dt) * 1e6) * stride;
        if (interval > kMaxSegyIntervalUs) invalid("SEG-Y rev1 sample interval must not exceed 32767 us.");
        seis3d_gather_info g{};
        g.spacing = c.spacing;
        g.frequency = c.frequency;
        g.source = *src;
        g.world_origin[0] = world_origin[0];
        g.world_origin[1] = world_origin[1];
        g.cmp_bin[0] = cmp_bin_size[0];
        g.cmp_bin[1] = cmp_bin_size[1];
        g.shot_id = shot_id;
        g.stride = stride;
        g.samples = samples;
        g.sample_interval_us = int(interval);
        g.first_sample_time = (stride - 0.5) * c.dt;
        const double h = c.spacing;
        double most = std::max(std::abs(world_origin[0] + src->x * h), std::abs(world_origin[1] + src->y * h));
        millimetres(src->z * h);
        for (int r = 0; r < count; ++r) {
            if (rec[r].x < 0 || rec[r].x >= c.nx || rec[r].y < 0 || rec[r].y >= c.ny || rec[r].z < 0 || rec[r].z >= c.nz)
                invalid("Seis3D: receiver outside the physical domain.");
            most = std::max({most, std::abs(world_origin[0] + rec[r].x * h), std::abs(world_origin[1] + rec[r].y * h)});
            millimetres(-rec[r].z * h);
            cmp_bin(g, rec[r]);
        }
        int factor = 1000;
        while (factor > 1 && most * factor > 2147483646.0) factor /= 10;
        if (!std::isfinite(most) || most * factor > 2147483646.0) invalid("World coordinates exceed the SEG-Y integer range.");
        g.coordinate_scalar = factor == 1 ? 1 : -factor;
        *out = g;
    });
}

int seis3d_export_gather(const seis3d_gather_info* info, const seis3d_receiver* rec, int count, const float* samples,
                         const char* directory, const char* prefix) {
    return guarded([&] {
        if (!info || !rec || !samples || !directory || count < 1 || count > kMaxReceivers || info->samples < 1 ||
            info->samples > kMaxSegySamples || info->sample_interval_us < 1 || info->sample_interval_us > kMaxSegyIntervalUs)
            invalid("SEG-Y rev1 sample count/interval exceeds its positive signed 16-bit range (32767).");
        const std::string base = shot_base(prefix, info->shot_id);
        const fs::path dir(directory);
        const char* endings[4] = {"_vx.sgy", "_vy.sgy", "_vz.sgy", "_geometry.csv"};
        for (const char* e : endings) {
            const fs::path p = dir / (base + e);
            if (fs::exists(p) || fs::exists(fs::path(p.string() + ".part")))
                invalid("Export files already exist; choose another prefix or directory.");
        }
        fs::create_directories(dir);
        for (int comp = 0; comp < 3; ++comp) write_component(*info, rec, count, samples, dir / (base + endings[comp]), comp);
        const fs::path csv_path = dir / (base + endings[3]);
        std::ofstream csv(csv_path);
        if (!csv) fail(SEIS3D_EHIP, "Cannot write geometry manifest.");
        csv << "shot,receiver,sx_m,sy_m,sz_depth_m,gx_m,gy_m,gz_depth_m,cmp_inline,cmp_crossline,first_sample_s,"
               "sample_interval_s,samples\n"
            << std::setprecision(12);
        const double h = info->spacing;
        for (int r = 0; r < count; ++r) {
            const auto b = cmp_bin(*info, rec[r]);
            csv << info->shot_id << ',' << r + 1 << ',' << info->world_origin[0] + info->source.x * h << ','
                << info->world_origin[1] + info->source.y * h << ',' << info->source.z * h << ','
                << info->world_origin[0] + rec[r].x * h << ',' << info->world_origin[1] + rec[r].y * h << ',' << rec[r].z * h
                << ',' << b.first << ',' << b.second << ',' << info->first_sample_time << ','
                << info->sample_interval_us * 1e-6 << ',' << info->samples << '\n';
        }
        csv.close();
        if (!csv) fail(SEIS3D_EHIP, "Cannot finalize geometry manifest.");
    });
}

int seis3d_segy_inspect(const char* path, seis3d_segy_info* info) {
    return guarded([&] {
        Reader r(path);
        r.read_all(nullptr, nullptr, nullptr, nullptr);   // validates every trace
        if (info) *info = r.info;
    });
}

int seis3d_segy_read(const char* path, float* samples, double* coordinates, double* first_times, int32_t* trace_ids) {
    return guarded([&] {
        Reader r(path);
        // Validate the whole file before touching any output.
        r.read_all(nullptr, nullptr, nullptr, nullptr);
        r.in.clear();
        r.read_all(samples, coordinates, first_times, trace_ids);
    });
}

int seis3d_s3dm_inspect(const char* path, seis3d_ctx* ctx) {
    return guarded([&] {
        if (!path || !ctx) invalid("Seis3D: invalid model request.");
        std::ifstream in(path, std::ios::binary);
        if (!in) invalid("Cannot open Seis3D model.");
        char magic[8]{};
        in.read(magic, 8);
        const std::uint32_t version = le32(in);
        if (std::memcmp(magic, "S3DMOD1", 8) != 0 || (version != 1 && version != 2)) invalid("Not a supported .s3dm model.");
        seis3d_ctx c = *ctx;
        c.nx = int(le32(in));
        c.ny = int(le32(in));
        c.nz = int(le32(in));
        c.pml = int(le32(in));
        const std::uint32_t boundary = le32(in);
        if (boundary > (version == 1 ? 1u : 2u)) invalid("Invalid model boundary flag.");
        c.free_surface = boundary == 1;
        c.reflecting = boundary == 2;
        const std::uint32_t sbits = le32(in);
        std::memcpy(&c.spacing, &sbits, 4);
        c.model_kind = SEIS3D_MODEL_IMPORTED;
        seis3d_ctx layout = c;
        layout.manual_dt = 0;
        layout.max_vp = 1000.0f;   // shape check only; the payload scan sets the real ceiling
        validate(layout);
        const std::size_t plane = std::size_t(c.nx) * c.ny * 3;
        std::error_code ec;
        if (fs::file_size(path, ec) != 36 + plane * std::size_t(c.nz) * sizeof(float) || ec)
            invalid("Model payload size does not match dimensions.");
        std::vector<float> buf(plane);
        float vmax = 0;
        for (int z = 0; z < c.nz; ++z) {
            in.read(reinterpret_cast<char*>(buf.data()), std::streamsize(plane * 4));
            if (!in) invalid("Cannot read model plane.");
            for (std::size_t i = 0; i < plane; i += 3) {
                if (!valid_material(buf[i], buf[i + 1], buf[i + 2])) invalid("Model contains an invalid material cell.");
                vmax = std::max(vmax, buf[i]);
            }
        }
        c.max_vp = vmax;
        validate(c);
        *ctx = c;
    });
}

int seis3d_s3dm_load(seis3d_t* t, const char* path) {
    return guarded([&] {
        require_session(t);
        seis3d_ctx c = t->c;
        if (seis3d_s3dm_inspect(path, &c) != SEIS3D_OK) fail(SEIS3D_EINVAL, g_error);
        if (c.nx != t->c.nx || c.ny != t->c.ny || c.nz != t->c.nz)
            invalid("Seis3D: the model's dimensions differ from the session's.");
        // The session's dt and CPML tables were built for its spacing and boundaries.
        if (std::memcmp(&c.spacing, &t->c.spacing, sizeof(float)) != 0 || c.pml != t->c.pml ||
            c.free_surface != t->c.free_surface || c.reflecting != t->c.reflecting)
            invalid("Seis3D: the model's spacing, CPML thickness or boundaries differ from the session's.");
        std::ifstream in(path, std::ios::binary);
        in.seekg(36);
        std::vector<float> plane(std::size_t(c.nx) * c.ny * 3);
        for (int z = 0; z < c.nz; ++z) {
            in.read(reinterpret_cast<char*>(plane.data()), std::streamsize(plane.size() * 4));
            if (!in) invalid("Seis3D model changed or became unreadable during import.");
            const int rc = seis3d_upload_plane(t, z, plane.data());
            if (rc != SEIS3D_OK) fail(rc, g_error);
        }
        const int rc = seis3d_finish_upload(t);
        if (rc != SEIS3D_OK) fail(rc, g_error);
    });
}

int seis3d_s3dm_save(seis3d_t* t, const char* path) {
    return guarded([&] {
        require_session(t);
        write_model(path, t->c, [&](int z, float* plane) {
            const int rc = seis3d_download_plane(t, z, plane);
            if (rc != SEIS3D_OK) fail(rc, g_error);
        });
    });
}

int seis3d_s3dm_write(const char* path, const seis3d_ctx* ctx, const float* vp_vs_rho) {
    return guarded([&] {
        if (!ctx || !vp_vs_rho) invalid("Seis3D: invalid model request.");
        const std::size_t plane = std::size_t(ctx->nx) * ctx->ny * 3;
        write_model(path, *ctx, [&](int z, float* out) {
            std::memcpy(out, vp_vs_rho + plane * std::size_t(z), plane * sizeof(float));
        });
    });
}

}  // extern "C"
