// seis3d_noalloc.hip - seis3d_extract_views_noalloc: the packed views of seis3d_extract_views
// staged in a fixed device array of this library instead of per-call allocations, for callers
// that cannot allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).
//
// The array `scratch` below holds the staging of up to kStageSamples view samples (2 floats
// each) and the 5 statistics words of the extraction kernel.  Each GPU the library runs on
// holds its own copy (16 MiB + 256 B) until the process ends.  Host code never passes the
// array's name to a kernel or a copy: the name is a host-side stand-in, not a device address.
// The device address is taken once per device with hipGetSymbolAddress and kept per device.
// Every call on one GPU shares that GPU's array, so calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <string>

#include "../../include/seis3d.h"
#include "seis3d_host.hpp"

namespace {

constexpr std::size_t kStageSamples = std::size_t(2) << 20;          // 2 Mi samples
constexpr std::size_t kStageFloats = 2 * kStageSamples;              // [wave, material] per sample
constexpr std::size_t kStatWords = 64;                               // 5 used, padded to 256 B
constexpr int kMaxDevices = 64;

static __device__ __align__(256) float scratch[kStageFloats + kStatWords];

std::mutex g_scratch_mutex;
float* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up on the first call there).
float* scratch_address(int device) {
    using seis3d::host::check;
    using seis3d::host::invalid;
    if (device < 0 || device >= kMaxDevices) invalid("Seis3D: device index outside the no-allocation table.");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[device]) {
        void* p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "locate no-allocation staging");
        if (!p) seis3d::host::fail(SEIS3D_EHIP, "Seis3D / locate no-allocation staging: null device address");
        g_scratch[device] = static_cast<float*>(p);
    }
    return g_scratch[device];
}

}  // namespace

extern "C" int seis3d_extract_views_noalloc(seis3d_t* t, float* out, size_t capacity, const int32_t slices[3],
                                            int component, int material_component, int volume, int include_boundary,
                                            seis3d_view_stats* stats) {
    using namespace seis3d::host;
    return guarded([&] {
        require_session(t);
        int device = -1;
        check(hipGetDevice(&device), "query current device");
        if (device != t->device)
            fail(SEIS3D_ELIFECYCLE, "Seis3D: make the session's device current before extracting views.");
        const long long count = view_samples(t->c, volume);
        if (count < 0 || std::size_t(count) > kStageSamples)
            invalid("Seis3D: the views need " + std::to_string(count) + " samples; the no-allocation staging holds " +
                    std::to_string(kStageSamples) + " (use seis3d_extract_views, or no volume).");
        float* base = scratch_address(device);
        extract_views_staged(t, out, capacity, slices, component, material_component, volume, include_boundary, stats,
                             base, reinterpret_cast<unsigned*>(base + kStageFloats));
    });
}
