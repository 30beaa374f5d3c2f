// seis2d_io.cpp - 2D shot records and shot packages (host only).
//
//  * writeSegyCaptureFile2D (Main_cpp/SeismicIO.cpp:296-414): one velocity
//    component as SEG-Y with IEEE big-endian samples and the acquisition trace
//    header fields of SeismicIO::TraceMetadata.
//  * readRecordSegy2D (src/Seis2D/RecordData2D.cpp:35-65): fixed-length
//    rev0/rev1 records in sample formats 1, 2, 3, 5 and 8.
//  * exportShotPackage2D (src/Seis2D/ShotExport2D.cpp:76-187): both records,
//    geometry and record CSVs, the frozen model (model.s2dm) and wavelet, and
//    a JSON manifest of the package with each file's SHA-256, published by
//    renaming a staging folder.
// The on-disk layouts (header fields, model layout, manifest schema) are the
// reference's so packages interchange; the SHA-256 is this file's own
// implementation of FIPS 180-4.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/seis2d.h"
#include "seis2d_host.hpp"

using namespace seis2d::host;
namespace fs = std::filesystem;

namespace {

constexpr float kSourceInjectionScale = 1e7f;   // SolverNumerics.h:12

[[noreturn]] void io_error(const std::string& m) { fail(SEIS2D_EIO, m); }

// ---- SHA-256 (FIPS 180-4) ----

// Round constants: the first 32 fraction bits of the cube roots of the first
// 64 primes; initial hash: of the square roots of the first 8 primes.  Both
// are derived exactly with integer roots of p * 2^96 and p * 2^64.
struct Sha256Constants {
    std::uint32_t k[64], h0[8];
    Sha256Constants() {
        int primes[64], n = 0;
        for (int p = 2; n < 64; ++p) {
            bool prime = true;
            for (int d = 2; d * d <= p; ++d) prime = prime && p % d != 0;
            if (prime) primes[n++] = p;
        }
        using u128 = unsigned __int128;
        for (int i = 0; i < 64; ++i) {
            const u128 target = u128(std::uint64_t(primes[i])) << 96;
            std::uint64_t lo = 0, hi = std::uint64_t(1) << 40;   // cbrt(311 * 2^96) < 2^35
            while (lo < hi) {
                const std::uint64_t mid = lo + (hi - lo + 1) / 2;
                if (u128(mid) * mid * mid <= target) lo = mid; else hi = mid - 1;
            }
            k[i] = std::uint32_t(lo);
        }
        for (int i = 0; i < 8; ++i) {
            const u128 target = u128(std::uint64_t(primes[i])) << 64;
            std::uint64_t lo = 0, hi = std::uint64_t(1) << 40;
            while (lo < hi) {
                const std::uint64_t mid = lo + (hi - lo + 1) / 2;
                if (u128(mid) * mid <= target) lo = mid; else hi = mid - 1;
            }
            h0[i] = std::uint32_t(lo);
        }
    }
};

class Sha256 {
public:
    Sha256() {
        static const Sha256Constants c;
        C = &c;
        std::memcpy(h, c.h0, sizeof h);
    }
    void update(const unsigned char* p, std::size_t n) {
        bits += std::uint64_t(n) * 8;
        while (n) {
            const std::size_t take = std::min(n, 64 - fill);
            std::memcpy(block + fill, p, take);
            fill += take; p += take; n -= take;
            if (fill == 64) { compress(); fill = 0; }
        }
    }
    std::string hex() {
        const std::uint64_t total = bits;
        const unsigned char one = 0x80, zero = 0;
        update(&one, 1);
        while (fill != 56) update(&zero, 1);
        unsigned char len[8];
        for (int i = 0; i < 8; ++i) len[i] = static_cast<unsigned char>(total >> (56 - 8 * i));
        update(len, 8);
        char out[65];
        for (int i = 0; i < 8; ++i) std::snprintf(out + 8 * i, 9, "%08x", h[i]);
        return std::string(out, 64);
    }

private:
    static std::uint32_t rotr(std::uint32_t x, int r) { return (x >> r) | (x << (32 - r)); }
    void compress() {
        std::uint32_t w[64];
        for (int i = 0; i < 16; ++i)
            w[i] = (std::uint32_t(block[4 * i]) << 24) | (std::uint32_t(block[4 * i + 1]) << 16) |
                   (std::uint32_t(block[4 * i + 2]) << 8) | block[4 * i + 3];
        for (int i = 16; i < 64; ++i) {
            const std::uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
            const std::uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
            w[i] = w[i - 16] + s0 + w[i - 7] + s1;
        }
        std::uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], k = h[7];
        for (int i = 0; i < 64; ++i) {
            const std::uint32_t t1 = k + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + C->k[i] + w[i];
            const std::uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
            k = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
        }
        h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += k;
    }
    const Sha256Constants* C;
    std::uint32_t h[8];
    unsigned char block[64];
    std::size_t fill = 0;
    std::uint64_t bits = 0;
};

std::string sha256_of(const fs::path& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) io_error("Cannot read package for SHA-256.");
    Sha256 s;
    std::vector<unsigned char> buf(1 << 16);
    while (in) {
        in.read(reinterpret_cast<char*>(buf.data()), std::streamsize(buf.size()));
        if (in.gcount() > 0) s.update(buf.data(), std::size_t(in.gcount()));
    }
    if (!in.eof()) io_error("Package file changed or became unreadable while hashing.");
    return s.hex();
}

// ---- big-endian fields ----

void put16(unsigned char* o, int at, int v) {
    o[at] = static_cast<unsigned char>(unsigned(v) >> 8);
    o[at + 1] = static_cast<unsigned char>(v);
}
void put32(unsigned char* o, int at, std::uint32_t v) {
    for (int k = 0; k < 4; ++k) o[at + k] = static_cast<unsigned char>(v >> (24 - 8 * k));
}
unsigned get16(const unsigned char* p) { return (unsigned(p[0]) << 8) | p[1]; }
int get16s(const unsigned char* p) {
    const unsigned v = get16(p);
    return v < 32768 ? int(v) : int(v) - 65536;
}
std::uint32_t get32(const unsigned char* p) {
    return (std::uint32_t(p[0]) << 24) | (std::uint32_t(p[1]) << 16) | (std::uint32_t(p[2]) << 8) | p[3];
}

void write_segy(const fs::path& path, const float* traces, int n, int ns, const seis2d_trace_meta* meta, float dt) {
    if (!traces || n < 1 || ns < 1 || !std::isfinite(dt) || dt <= 0)
        invalid("SEG-Y export requires samples and a positive interval.");
    const long long interval = std::llround(double(dt) * 1000000);
    if (ns > 32767 || interval < 1 || interval > 32767 || !meta)
        invalid("SEG-Y sample count/interval must fit 1..32767 and geometry must match.");
    for (long long i = 0; i < static_cast<long long>(n) * ns; ++i)
        if (!std::isfinite(traces[i])) invalid("SEG-Y export contains invalid amplitudes.");
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out) io_error("Cannot create SEG-Y capture file.");
    std::array<char, 3200> text;
    text.fill(' ');
    const char line[] = "C01 2D ELASTIC SHOT GATHER, PARTICLE VELOCITY M/S, IEEE FLOAT32 BIG-ENDIAN";
    std::memcpy(text.data(), line, sizeof(line) - 1);
    out.write(text.data(), text.size());
    std::array<unsigned char, 400> bin{};
    put16(bin.data(), 16, int(interval));
    put16(bin.data(), 20, ns);
    put16(bin.data(), 24, 5);
    out.write(reinterpret_cast<const char*>(bin.data()), bin.size());
    std::vector<unsigned char> data(std::size_t(ns) * 4);
    for (int r = 0; r < n; ++r) {
        std::array<unsigned char, 240> t{};
        const seis2d_trace_meta& m = meta[r];
        put32(t.data(), 0, std::uint32_t(r + 1));
        put32(t.data(), 8, std::uint32_t(m.field_record));
        put32(t.data(), 12, std::uint32_t(m.trace_in_record));
        put32(t.data(), 72, std::uint32_t(m.source_x));
        put32(t.data(), 76, std::uint32_t(m.source_y));
        put32(t.data(), 80, std::uint32_t(m.group_x));
        put32(t.data(), 84, std::uint32_t(m.group_y));
        put16(t.data(), 70, m.coordinate_scalar);
        put16(t.data(), 88, 1);
        put16(t.data(), 114, ns);
        put16(t.data(), 116, int(interval));
        put16(t.data(), 202, 6);   // trace values in m/s
        for (int k = 0; k < ns; ++k) {
            std::uint32_t bits;
            std::memcpy(&bits, traces + std::size_t(r) * ns + k, 4);
            put32(data.data(), 4 * k, bits);
        }
        out.write(reinterpret_cast<const char*>(t.data()), t.size());
        out.write(reinterpret_cast<const char*>(data.data()), std::streamsize(data.size()));
    }
    out.close();
    if (!out) io_error("Cannot finalize SEG-Y capture file; check disk space.");
}

struct Record {
    seis2d_record_info info{};
    std::uint64_t offset = 0;
    int bytes = 4;
};

void exact(std::istream& in, void* p, std::size_t n) {
    if (!in.read(static_cast<char*>(p), std::streamsize(n))) io_error("SEG-Y input is truncated.");
}

float decode(const unsigned char* p, int format) {
    double v = 0;
    if (format == 5) {
        const std::uint32_t b = get32(p);
        float f;
        std::memcpy(&f, &b, 4);
        v = f;
    } else if (format == 1) {
        const std::uint32_t b = get32(p);
        v = std::ldexp(double(b & 0xffffff) / 16777216.0, 4 * (int((b >> 24) & 127) - 64));
        if (b >> 31) v = -v;
    } else if (format == 2) {
        const std::uint32_t b = get32(p);
        std::int32_t s;
        std::memcpy(&s, &b, 4);
        v = s;
    } else if (format == 3) {
        v = get16s(p);
    } else {
        v = p[0] < 128 ? int(p[0]) : int(p[0]) - 256;
    }
    if (!std::isfinite(v) || std::abs(v) > std::numeric_limits<float>::max())
        io_error("SEG-Y contains invalid amplitudes.");
    return float(v);
}

// Reads (and fully validates) a record; outputs may be null.
Record read_record(const char* path, float* traces, double* times) {
    if (!path) invalid("Seis2D: null path.");
    std::ifstream in(path, std::ios::binary);
    if (!in) io_error("Cannot open SEG-Y record.");
    std::error_code ec;
    const std::uint64_t size = fs::file_size(path, ec);
    if (ec || size < 3840) io_error("SEG-Y input is truncated.");
    std::array<unsigned char, 400> bin{};
    in.seekg(3200);
    exact(in, bin.data(), bin.size());
    Record r;
    const int format = int(get16(bin.data() + 24)), ns = int(get16(bin.data() + 20)), dt = int(get16(bin.data() + 16));
    const int extended = get16s(bin.data() + 304);
    if (get16(bin.data() + 300) > 256 || extended < 0) io_error("Unsupported SEG-Y revision or extended header count.");
    if (format != 1 && format != 2 && format != 3 && format != 5 && format != 8) io_error("Unsupported SEG-Y sample format.");
    r.bytes = format == 3 ? 2 : format == 8 ? 1 : 4;
    if (ns < 1 || ns > 32767 || dt < 1) io_error("Invalid SEG-Y sample count or interval.");
    r.offset = 3600 + std::uint64_t(extended) * 3200;
    const std::uint64_t stride = 240 + std::uint64_t(ns) * r.bytes;
    if (r.offset >= size || (size - r.offset) % stride) io_error("SEG-Y trace lengths do not match the file size.");
    const std::uint64_t channels = (size - r.offset) / stride;
    if (channels < 1 || channels > 8192 || channels * std::uint64_t(ns) * sizeof(float) > 512ull * 1024 * 1024)
        io_error("SEG-Y record exceeds the 8192-trace or 512 MiB viewer budget.");
    r.info.traces = int(channels);
    r.info.samples = ns;
    r.info.sample_interval_us = dt;
    r.info.format = format;
    in.seekg(std::streamoff(r.offset));
    std::vector<unsigned char> payload(std::size_t(ns) * r.bytes);
    bool velocity = true;
    for (std::uint64_t c = 0; c < channels; ++c) {
        std::array<unsigned char, 240> h{};
        exact(in, h.data(), h.size());
        exact(in, payload.data(), payload.size());
        const int tns = int(get16(h.data() + 114)), tdt = int(get16(h.data() + 116));
        if ((tns && tns != ns) || (tdt && tdt != dt)) io_error("Variable SEG-Y trace lengths or intervals are not supported.");
        const int scalar = get16s(h.data() + 214);
        const double factor = scalar == 0 ? 1.0 : scalar > 0 ? double(scalar) : 1.0 / -double(scalar);
        if (times) times[c] = get16s(h.data() + 108) * factor * .001;
        velocity = velocity && get16(h.data() + 202) == 6;
        for (int i = 0; i < ns; ++i) {
            const float v = decode(payload.data() + std::size_t(i) * r.bytes, format);
            if (traces) traces[c * std::uint64_t(ns) + std::uint64_t(i)] = v;
        }
    }
    r.info.velocity_units = velocity ? 1 : 0;
    return r;
}

void validate_prefix(const std::string& prefix) {
    if (prefix.empty() || prefix.size() > 120 || prefix == "." || prefix == ".." || prefix.back() == '.' ||
        prefix.back() == ' ')
        invalid("Shot prefix must be 1-120 UTF-8 bytes without a trailing dot or space.");
    for (unsigned char c : prefix)
        if (c < 32 || std::strchr("<>:\"/\\|?*", c))
            invalid("Shot prefix must be a filename, without path separators or reserved characters.");
    std::string stem = prefix.substr(0, prefix.find('.'));
    for (char& c : stem)
        if (c >= 'a' && c <= 'z') c = char(c - 32);
    if (stem == "CON" || stem == "PRN" || stem == "AUX" || stem == "NUL" ||
        (stem.size() == 4 && (stem.substr(0, 3) == "COM" || stem.substr(0, 3) == "LPT") && stem[3] >= '1' && stem[3] <= '9'))
        invalid("Shot prefix is a reserved Windows filename.");
}

std::string json_text(const char* v) {
    std::string out = "\"";
    for (const char* p = v ? v : ""; *p; ++p) {
        const unsigned char c = static_cast<unsigned char>(*p);
        if (c == '"' || c == '\\') {
            out += '\\';
            out += char(c);
        } else if (c < 32) {
            char buf[8];
            std::snprintf(buf, sizeof buf, "\\u%04x", c);
            out += buf;
        } else {
            out += char(c);
        }
    }
    return out + "\"";
}

// Shortest form that reads back exactly (as the reference's max_digits10 streams).
std::string num(double v) {
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}

void write_words(std::ostream& out, const float* v, std::size_t n) {
    for (std::size_t i = 0; i < n; ++i) {
        std::uint32_t b;
        std::memcpy(&b, v + i, 4);
        const unsigned char w[4] = {static_cast<unsigned char>(b), static_cast<unsigned char>(b >> 8),
                                    static_cast<unsigned char>(b >> 16), static_cast<unsigned char>(b >> 24)};
        out.write(reinterpret_cast<const char*>(w), 4);
    }
}
void write_u32le(std::ostream& out, std::uint32_t b) {
    const unsigned char w[4] = {static_cast<unsigned char>(b), static_cast<unsigned char>(b >> 8),
                                static_cast<unsigned char>(b >> 16), static_cast<unsigned char>(b >> 24)};
    out.write(reinterpret_cast<const char*>(w), 4);
}

std::ofstream open_out(const fs::path& p, bool binary) {
    std::ofstream out(p, binary ? std::ios::binary | std::ios::trunc : std::ios::trunc);
    if (!out) io_error("Cannot create package file " + p.filename().string());
    return out;
}
void close_out(std::ofstream& out) {
    out.close();
    if (!out) io_error("Package write failed; check disk space.");
}

// validatePackage (ShotExport2D.cpp:76-100).
void validate_package(const seis2d_export_request& r, const seis2d_frozen_model& m, const float* vx, const float* vz,
                      int traces, const seis2d_trace_meta* geometry) {
    validate_prefix(r.prefix ? r.prefix : "");
    if (!r.directory || !*r.directory || r.shot < 1 || r.samples < 1 || r.samples > 32767 ||
        !std::isfinite(r.sample_interval) || r.sample_interval <= 0 || r.sample_interval * 1e6 < .5 ||
        r.sample_interval * 1e6 >= 32767.5)
        invalid("Choose an output directory, valid shot ID and SEG-Y-compatible samples/interval.");
    if (traces < 1 || !vx || !vz || !geometry) invalid("Shot components and geometry differ.");
    for (const float* c : {vx, vz})
        for (long long i = 0; i < static_cast<long long>(traces) * r.samples; ++i)
            if (!std::isfinite(c[i])) invalid("Shot contains invalid amplitudes.");
    if (m.nx < 1 || m.nz < 1 || m.npml < 0 || m.formulation < 1 || m.formulation > 3 || m.sampling_stride < 1 ||
        m.capture_start_step < 0 || m.source_type < 0 || m.source_type > 4 || m.source_x < 0 || m.source_x >= m.nx ||
        m.source_z < 0 || m.source_z >= m.nz || !std::isfinite(m.spacing) || m.spacing <= 0 ||
        !std::isfinite(m.solver_dt) || m.solver_dt <= 0 || !std::isfinite(m.frequency) || m.frequency <= 0 ||
        !std::isfinite(m.delay) || !std::isfinite(m.source_amplitude) || !std::isfinite(m.source_angle))
        invalid("Frozen model parameters are invalid.");
    if (r.include_model) {
        const std::uint64_t count = std::uint64_t(m.nx) * std::uint64_t(m.nz);
        if (count * 12 > 256ull * 1024 * 1024 || !m.lambda || !m.mu || !m.rho || !m.wavelet || m.wavelet_count < 1)
            invalid("Frozen model arrays are absent, inconsistent or exceed 256 MiB.");
        for (const float* a : {m.lambda, m.mu, m.rho})
            for (std::uint64_t i = 0; i < count; ++i)
                if (!std::isfinite(a[i])) invalid("Frozen model contains non-finite values.");
        for (int i = 0; i < m.wavelet_count; ++i)
            if (!std::isfinite(m.wavelet[i])) invalid("Frozen model contains non-finite values.");
    }
}

}  // namespace

extern "C" {

int seis2d_write_segy(const char* path, const float* traces, int n_traces, int n_samples, const seis2d_trace_meta* meta,
                      float dt) {
    return guarded([&] {
        if (!path) invalid("Seis2D: null path.");
        write_segy(path, traces, n_traces, n_samples, meta, dt);
    });
}

int seis2d_read_segy_info(const char* path, seis2d_record_info* info) {
    return guarded([&] {
        const Record r = read_record(path, nullptr, nullptr);
        if (info) *info = r.info;
    });
}

int seis2d_read_segy(const char* path, float* traces, double* first_times) {
    return guarded([&] {
        read_record(path, nullptr, nullptr);   // validate everything before writing outputs
        read_record(path, traces, first_times);
    });
}

int seis2d_validate_prefix(const char* prefix) {
    return guarded([&] { validate_prefix(prefix ? prefix : ""); });
}

int seis2d_sha256_file(const char* path, char out_hex[65]) {
    return guarded([&] {
        if (!path || !out_hex) invalid("Seis2D: null argument.");
        const std::string h = sha256_of(path);
        std::memcpy(out_hex, h.c_str(), 65);
    });
}

int seis2d_export_shot_package(const seis2d_export_request* request, const seis2d_frozen_model* model, const float* vx,
                               const float* vz, int traces, const seis2d_trace_meta* geometry, char* out_path,
                               size_t out_capacity) {
    return guarded([&] {
        if (!request || !model) invalid("Seis2D: null export request.");
        const seis2d_export_request& r = *request;
        const seis2d_frozen_model& m = *model;
        validate_package(r, m, vx, vz, traces, geometry);
        char name[160];
        std::snprintf(name, sizeof name, "%s_%04d", r.prefix, r.shot);
        const fs::path base = fs::absolute(r.directory).lexically_normal();
        const fs::path final_dir = base / name;
        if (fs::exists(final_dir)) io_error("Shot destination exists. Choose a new name or session.");
        fs::create_directories(base);
        const auto serial = std::chrono::steady_clock::now().time_since_epoch().count();
        const fs::path staging = base / (std::string(name) + ".part-" + std::to_string(serial));
        if (!fs::create_directory(staging)) io_error("Cannot create unique shot staging directory.");
        try {
            const int ns = r.samples;
            write_segy(staging / "record_vx.sgy", vx, traces, ns, geometry, r.sample_interval);
            write_segy(staging / "record_vz.sgy", vz, traces, ns, geometry, r.sample_interval);
            std::vector<std::string> files = {"record_vx.sgy", "record_vz.sgy", "geometry.csv"};
            {
                auto out = open_out(staging / "geometry.csv", false);
                out << "trace,source_x_m,source_z_m,receiver_x_m,receiver_z_m\n";
                for (int i = 0; i < traces; ++i) {
                    const seis2d_trace_meta& g = geometry[i];
                    const double s = g.coordinate_scalar < 0 ? 1.0 / -double(g.coordinate_scalar)
                                     : g.coordinate_scalar == 0 ? 1.0 : double(g.coordinate_scalar);
                    out << g.trace_in_record << ',' << num(g.source_x * s) << ',' << num(g.source_y * s) << ','
                        << num(g.group_x * s) << ',' << num(g.group_y * s) << '\n';
                }
                close_out(out);
            }
            if (r.include_csv) {
                auto out = open_out(staging / "record_vz.csv", false);
                out << "time_s";
                for (int c = 0; c < traces; ++c) out << ",trace_" << c + 1;
                out << '\n';
                for (int i = 0; i < ns; ++i) {
                    out << num(double(i) * r.sample_interval);
                    for (int c = 0; c < traces; ++c) out << ',' << num(vz[std::size_t(c) * ns + i]);
                    out << '\n';
                }
                close_out(out);
                files.push_back("record_vz.csv");
            }
            if (r.include_model) {
                auto out = open_out(staging / "model.s2dm", true);
                out.write("S2DMOD1\0", 8);
                write_u32le(out, std::uint32_t(m.nx));
                write_u32le(out, std::uint32_t(m.nz));
                std::uint32_t bits;
                std::memcpy(&bits, &m.spacing, 4);
                write_u32le(out, bits);
                const std::size_t count = std::size_t(m.nx) * std::size_t(m.nz);
                write_words(out, m.lambda, count);
                write_words(out, m.mu, count);
                write_words(out, m.rho, count);
                close_out(out);
                auto wave = open_out(staging / "source_wavelet.f32le", true);
                write_words(wave, m.wavelet, std::size_t(m.wavelet_count));
                close_out(wave);
                files.push_back("model.s2dm");
                files.push_back("source_wavelet.f32le");
            }
            const char* mechanisms[] = {"vertical body force", "horizontal body force", "isotropic stress", "shear stress",
                                        "rotated body force"};
            const auto now_us = std::chrono::duration_cast<std::chrono::microseconds>(
                                    std::chrono::system_clock::now().time_since_epoch()).count();
            std::ostringstream j;
            j << "{\n  \"schema\": \"SeisWaveShot2D-1\",\n  \"software\": \"seiswave-mi355x\",\n  \"version\": "
              << json_text(r.version) << ",\n  \"gpu\": " << json_text(r.device) << ",\n  \"shot_id\": " << r.shot
              << ",\n  \"created_unix_us\": " << now_us << ",\n  \"traces\": " << traces << ",\n  \"samples\": " << ns
              << ",\n  \"sample_interval_s\": " << num(r.sample_interval)
              << ",\n  \"segy_interval_us\": " << std::llround(double(r.sample_interval) * 1e6)
              << ",\n  \"units\": \"m/s\",\n  \"amplitudes\": \"raw; no display gain or AGC\",\n"
              << "  \"time_origin\": \"local acquisition start; SEG-Y delay is zero\",\n"
              << "  \"coordinates\": \"X horizontal; Z down from model grid origin; legacy SEG-Y Y stores Z\",\n"
              << "  \"model_layout\": "
              << json_text("S2DMOD1\\0 + uint32le nx,nz + float32le h; lambda,mu,rho planes; index=z*nx+x") << ",\n"
              << "  \"model_included\": " << (r.include_model ? "true" : "false") << ",\n  \"grid\": {\"nx\": " << m.nx
              << ", \"nz\": " << m.nz << ", \"spacing_m\": " << num(m.spacing) << ", \"npml\": " << m.npml
              << ", \"formulation\": " << m.formulation << ", \"free_surface\": " << (m.free_surface ? "true" : "false")
              << "},\n  \"solver\": {\"dt_s\": " << num(m.solver_dt) << ", \"sampling_stride\": " << m.sampling_stride
              << ", \"capture_start_step\": " << m.capture_start_step << ", \"nt\": " << m.wavelet_count
              << ", \"spatial_order\": 8, \"fast_math_build\": false},\n  \"source\": {\"x\": " << m.source_x
              << ", \"z\": " << m.source_z << ", \"type\": " << m.source_type
              << ", \"mechanism\": " << json_text(mechanisms[m.source_type])
              << ", \"numerical_injection_scale\": " << num(kSourceInjectionScale)
              << ", \"amplitude\": " << num(m.source_amplitude) << ", \"angle_deg\": " << num(m.source_angle)
              << ", \"f0_hz\": " << num(m.frequency) << ", \"delay_s\": " << num(m.delay)
              << ", \"multiple_sources\": " << (m.multiple_sources ? "true" : "false") << "},\n"
              << "  \"reproduction_note\": \"Continuation captures lack the initial dynamic wavefield; multi-source "
                 "events are not archived. This package alone is not a replay checkpoint.\",\n  \"files\": [\n";
            for (std::size_t i = 0; i < files.size(); ++i)
                j << "    {\"path\": " << json_text(files[i].c_str()) << ", \"bytes\": " << fs::file_size(staging / files[i])
                  << ", \"sha256\": " << json_text(sha256_of(staging / files[i]).c_str()) << '}'
                  << (i + 1 < files.size() ? "," : "") << '\n';
            j << "  ]\n}\n";
            {
                auto out = open_out(staging / "manifest.json", false);
                out << j.str();
                close_out(out);
            }
            if (fs::exists(final_dir)) io_error("Shot destination was created by another writer.");
            fs::rename(staging, final_dir);
        } catch (...) {
            std::error_code ignored;
            fs::remove_all(staging, ignored);
            throw;
        }
        if (out_path) {
            const std::string p = final_dir.string();
            if (out_capacity < p.size() + 1) invalid("Seis2D: output path buffer too small (package was written).");
            std::memcpy(out_path, p.c_str(), p.size() + 1);
        }
    });
}

}  // extern "C"
