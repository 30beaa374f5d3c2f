// seis3d_stream.hip - HBM bandwidth probes: the measured ceiling the step kernels are
// compared with.
//
// Every probe streams whole arrays with 16-byte accesses per lane, the loads and stores
// non-temporal (streamed data is not kept in the caches), and the grid is timed at 2, 4, 8, 16, 32
// and 64 workgroups per CU with the fastest reported: the ceiling is the best this hardware does
// with the pattern.  (Cached accesses and at least 4 workgroups per CU measured 12-20 % lower.)
// Kinds:
//   read   - sum of every float of a, grid-stride, eight requests in flight per lane (one float per
//            thread written back, so the loads stay);
//   write  - fill b, grid-stride;
//   copy   - b = a through registers, each workgroup copying contiguous runs of 8 x 256 float4s;
//   dma    - b = a through LDS: global_load_lds_dwordx4 stages a wave's 1 KiB, the wave
//            waits for it, reads it back and stores it (the staging the step kernels use);
//   mix10_3 / mix12_6 - out_m = sum_k c_mk in_k over 10 inputs and 3 outputs (the velocity
//            half-step's streams) or 12 and 6 (the stress half-step's), elementwise.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "../../include/seis3d.h"
#include "seis3d_host.hpp"

namespace seis3d {
namespace {

constexpr int kBlock = 256;
constexpr int kUnroll = 4;
constexpr int kReadUnroll = 8;
constexpr int kCopyRun = 8;

// Clang's native four-float vector (the non-temporal builtins take vectors of scalars).
typedef float v4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4 load_nt(const v4* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ void store_nt(v4 v, v4* p) { __builtin_nontemporal_store(v, p); }

__global__ __launch_bounds__(kBlock) void read_probe(const v4* __restrict__ a, std::int64_t n4, float* out) {
    const std::int64_t stride = std::int64_t(gridDim.x) * kBlock;
    std::int64_t i = std::int64_t(blockIdx.x) * kBlock + threadIdx.x;
    v4 acc = {0, 0, 0, 0};
    for (; i + (kReadUnroll - 1) * stride < n4; i += kReadUnroll * stride) {
        v4 v[kReadUnroll];
#pragma unroll
        for (int u = 0; u < kReadUnroll; ++u) v[u] = load_nt(a + i + u * stride);
#pragma unroll
        for (int u = 0; u < kReadUnroll; ++u) acc += v[u];
    }
    for (; i < n4; i += stride) acc += load_nt(a + i);
    out[std::int64_t(blockIdx.x) * kBlock + threadIdx.x] = acc.x + acc.y + acc.z + acc.w;
}

__global__ __launch_bounds__(kBlock) void write_probe(v4* __restrict__ b, std::int64_t n4, float value) {
    const std::int64_t stride = std::int64_t(gridDim.x) * kBlock;
    const v4 v = {value, value, value, value};
    for (std::int64_t i = std::int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n4; i += stride) store_nt(v, b + i);
}

// Each workgroup copies runs of kCopyRun * kBlock consecutive float4s: all of a run's loads are in
// flight before its stores.
__global__ __launch_bounds__(kBlock) void copy_probe(const v4* __restrict__ a, v4* __restrict__ b, std::int64_t n4) {
    constexpr std::int64_t run = std::int64_t(kCopyRun) * kBlock;
    for (std::int64_t base = std::int64_t(blockIdx.x) * run; base < n4; base += std::int64_t(gridDim.x) * run) {
        v4 v[kCopyRun];
#pragma unroll
        for (int u = 0; u < kCopyRun; ++u) {
            const std::int64_t i = base + u * kBlock + threadIdx.x;
            v[u] = i < n4 ? load_nt(a + i) : v4{0, 0, 0, 0};
        }
#pragma unroll
        for (int u = 0; u < kCopyRun; ++u) {
            const std::int64_t i = base + u * kBlock + threadIdx.x;
            if (i < n4) store_nt(v[u], b + i);
        }
    }
}

// Each wave owns a 1 KiB LDS slot per unrolled request.
__global__ __launch_bounds__(kBlock) void dma_probe(const v4* __restrict__ a, v4* __restrict__ b, std::int64_t n4) {
    __shared__ v4 slot[kUnroll][kBlock];
    const std::int64_t stride = std::int64_t(gridDim.x) * kBlock;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    std::int64_t i = std::int64_t(blockIdx.x) * kBlock + threadIdx.x;
    // Whole groups of kUnroll requests only: the tail is copied through registers, so every
    // LDS-DMA instruction of a wave reads inside the array.
    const std::int64_t whole = n4 / (kUnroll * stride) * (kUnroll * stride);
    for (; i < whole; i += kUnroll * stride) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
            __builtin_amdgcn_global_load_lds(static_cast<const void*>(a + i + u * stride), &slot[u][wave * 64], 16, 0, 0);
        __builtin_amdgcn_s_waitcnt(0x0070 | 0x0f00);   // vmcnt(0)
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) store_nt(slot[u][wave * 64 + lane], b + i + u * stride);
        __builtin_amdgcn_wave_barrier();
    }
    for (; i < n4; i += stride) store_nt(load_nt(a + i), b + i);
}

template <int IN, int OUT>
__global__ __launch_bounds__(kBlock) void mix_probe(const float* const* __restrict__ in, float* const* __restrict__ out,
                                                   std::int64_t n4) {
    const std::int64_t stride = std::int64_t(gridDim.x) * kBlock;
    const v4* src[IN];
    v4* dst[OUT];
#pragma unroll
    for (int k = 0; k < IN; ++k) src[k] = reinterpret_cast<const v4*>(in[k]);
#pragma unroll
    for (int m = 0; m < OUT; ++m) dst[m] = reinterpret_cast<v4*>(out[m]);
    for (std::int64_t i = std::int64_t(blockIdx.x) * kBlock + threadIdx.x; i < n4; i += stride) {
        v4 v[IN];
#pragma unroll
        for (int k = 0; k < IN; ++k) v[k] = load_nt(src[k] + i);
#pragma unroll
        for (int m = 0; m < OUT; ++m) {
            v4 r = {0, 0, 0, 0};
#pragma unroll
            for (int k = 0; k < IN; ++k) r += (1.0f / float(1 + k + m)) * v[k];
            store_nt(r, dst[m] + i);
        }
    }
}

}  // namespace
}  // namespace seis3d

using seis3d::host::check;
using seis3d::host::guarded;
using seis3d::host::invalid;

extern "C" int seis3d_bandwidth_probe(int kind, uint64_t bytes, int reps, double* gbs, double* us_per_launch) {
    return guarded([&] {
        if (kind < 0 || kind > 5 || reps < 1 || reps > 1000 || bytes < (1u << 20) || bytes > (uint64_t(64) << 30) ||
            !gbs)
            invalid("Seis3D: invalid bandwidth probe request.");
        using namespace seis3d;
        const int streams_in = kind == 4 ? 10 : kind == 5 ? 12 : 1;
        const int streams_out = kind == 4 ? 3 : kind == 5 ? 6 : 1;
        // `bytes` is the size of one array; every array is a whole number of float4s.
        const std::int64_t n4 = std::int64_t(bytes / 16);
        const std::size_t each = std::size_t(n4) * 16;
        int dev = 0, cus = 0;
        check(hipGetDevice(&dev), "query device");
        check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "query device");
        const int max_grid = std::max(1, cus) * 64;
        std::vector<float*> arrays(std::size_t(streams_in + streams_out), nullptr);
        float* partial = nullptr;
        float** table = nullptr;
        hipStream_t s = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        auto release = [&] {
            if (s) (void)hipStreamSynchronize(s);
            for (auto p : arrays) (void)hipFree(p);
            (void)hipFree(partial);
            (void)hipFree(table);
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (s) (void)hipStreamDestroy(s);
        };
        try {
            check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "create stream");
            check(hipEventCreate(&e0), "create event");
            check(hipEventCreate(&e1), "create event");
            for (auto& p : arrays) {
                check(hipMalloc(&p, each), "allocate probe array");
                check(hipMemsetAsync(p, 0, each, s), "clear probe array");
            }
            check(hipMalloc(&partial, std::size_t(max_grid) * kBlock * sizeof(float)), "allocate probe array");
            check(hipMalloc(&table, arrays.size() * sizeof(float*)), "allocate probe table");
            check(hipMemcpyAsync(table, arrays.data(), arrays.size() * sizeof(float*), hipMemcpyHostToDevice, s),
                  "upload probe table");
            auto launch = [&](int grid) {
                const auto* a = reinterpret_cast<const v4*>(arrays[0]);
                auto* b = reinterpret_cast<v4*>(arrays.back());
                switch (kind) {
                case 0: hipLaunchKernelGGL(read_probe, dim3(grid), dim3(kBlock), 0, s, a, n4, partial); break;
                case 1: hipLaunchKernelGGL(write_probe, dim3(grid), dim3(kBlock), 0, s, b, n4, 1.0f); break;
                case 2: hipLaunchKernelGGL(copy_probe, dim3(grid), dim3(kBlock), 0, s, a, b, n4); break;
                case 3: hipLaunchKernelGGL(dma_probe, dim3(grid), dim3(kBlock), 0, s, a, b, n4); break;
                case 4:
                    hipLaunchKernelGGL((mix_probe<10, 3>), dim3(grid), dim3(kBlock), 0, s, table, table + 10, n4);
                    break;
                default:
                    hipLaunchKernelGGL((mix_probe<12, 6>), dim3(grid), dim3(kBlock), 0, s, table, table + 12, n4);
                    break;
                }
                check(hipGetLastError(), "kernel launch");
            };
            double best = 0;
            for (int per_cu : {2, 4, 8, 16, 32, 64}) {
                const int grid = std::max(1, cus) * per_cu;
                launch(grid);   // warm-up
                check(hipEventRecord(e0, s), "record event");
                for (int r = 0; r < reps; ++r) launch(grid);
                check(hipEventRecord(e1, s), "record event");
                check(hipEventSynchronize(e1), "bandwidth probe");
                float ms = 0;
                check(hipEventElapsedTime(&ms, e0, e1), "read event");
                const double per = double(ms) * 1e-3 / reps;
                if (best == 0 || per < best) best = per;
            }
            const int touched = kind == 0 || kind == 1 ? 1 : streams_in + streams_out;
            *gbs = double(each) * touched / best * 1e-9;
            if (us_per_launch) *us_per_launch = best * 1e6;
        } catch (...) {
            release();
            throw;
        }
        release();
    });
}
