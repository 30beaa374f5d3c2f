// seis2d_kernels.hip - CDNA4 (gfx950) kernels of the 2D P-SV O(8)
// staggered-grid velocity-stress step: split-field PML (type 1), hybrid
// PML/FDM (type 2) and free surface (type 3), sources, the zero frame and
// receivers (restating CUDA_Kernels/Cuda_Check.cu:225-667,793-834).
//
// One thread per cell, 64 x 4 workgroups so that a wavefront covers 64
// consecutive x.  Expressions keep the reference's operand order; built with
// -ffp-contract=off and IEEE division, the results equal the FP32 oracle's.
// The reference build uses fast-math (approximate division), which this
// build does not reproduce.
//
// A step is three launches (four with the type-3 ghost mirror): the stress
// and velocity half-steps, then one workgroup that injects the sources,
// zeroes the frame and takes the receiver samples, in that order.
#include "seis2d_device.hpp"

namespace seis2d {

namespace {

constexpr int kR = 4;

__device__ __forceinline__ float sdiff(const float* __restrict__ a, const float* __restrict__ b, int i, int j) {
    return a[i] + b[i] - a[j] - b[j];
}
// Staggered derivative forward of k (taps k+s..k+4s minus k..k-3s) and backward (k..k+3s minus k-s..k-4s).
__device__ __forceinline__ float fwd(const Dev2D& g, const float* a, const float* b, int k, int s) {
    return g.c1 * sdiff(a, b, k + s, k) + g.c2 * sdiff(a, b, k + 2 * s, k - s) +
           g.c3 * sdiff(a, b, k + 3 * s, k - 2 * s) + g.c4 * sdiff(a, b, k + 4 * s, k - 3 * s);
}
__device__ __forceinline__ float bwd(const Dev2D& g, const float* a, const float* b, int k, int s) {
    return g.c1 * sdiff(a, b, k, k - s) + g.c2 * sdiff(a, b, k + s, k - 2 * s) +
           g.c3 * sdiff(a, b, k + 2 * s, k - 3 * s) + g.c4 * sdiff(a, b, k + 3 * s, k - 4 * s);
}
__device__ __forceinline__ float fwd1(const Dev2D& g, const float* __restrict__ a, int k, int s) {
    return g.c1 * (a[k + s] - a[k]) + g.c2 * (a[k + 2 * s] - a[k - s]) + g.c3 * (a[k + 3 * s] - a[k - 2 * s]) +
           g.c4 * (a[k + 4 * s] - a[k - 3 * s]);
}
__device__ __forceinline__ float bwd1(const Dev2D& g, const float* __restrict__ a, int k, int s) {
    return g.c1 * (a[k] - a[k - s]) + g.c2 * (a[k + s] - a[k - 2 * s]) + g.c3 * (a[k + 2 * s] - a[k - 3 * s]) +
           g.c4 * (a[k + 3 * s] - a[k - 4 * s]);
}
// Split-PML update (old (1/dt - d) + source) / (1/dt + d).
__device__ __forceinline__ float pml(float old, float inv, float d, float src) { return (old * (inv - d) + src) / (inv + d); }

__device__ __forceinline__ bool cell(const Dev2D& g, int& i, int& j, int row0) {
    j = int(blockIdx.x * blockDim.x + threadIdx.x);
    i = int(blockIdx.y * blockDim.y + threadIdx.y);
    return i >= row0 && i < g.nz - kR && j >= kR && j < g.nx - kR;
}

__device__ __forceinline__ void count_step(const Dev2D& g) {
    if (g.counter && blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) *g.counter += 1;
}

// updateType1StressKernel (:225-266) and, with `fsz`, the interior rows of
// updateType3StressKernel (:465-485), whose z damping is off near the surface.
__device__ __forceinline__ void split_stress(const Dev2D& g, int j, int k, float dzv, float dzhv) {
    float* const* f = g.f;
    const float dvx_dx = fwd(g, f[0], f[1], k, 1), dvz_dz = bwd(g, f[2], f[3], k, g.nx);
    const float l2 = 0.5f * (g.l2m[k] + g.l2m[k + 1]), la = 0.5f * (g.lambda[k] + g.lambda[k + 1]);
    const float dxh = g.dxh[j];
    f[4][k] = pml(f[4][k], g.inv_dt, dxh, l2 * dvx_dx);
    f[5][k] = pml(f[5][k], g.inv_dt, dzv, la * dvz_dz);
    f[6][k] = pml(f[6][k], g.inv_dt, dxh, la * dvx_dx);
    f[7][k] = pml(f[7][k], g.inv_dt, dzv, l2 * dvz_dz);
    const float dvz_dx = bwd(g, f[2], f[3], k, 1), dvx_dz = fwd(g, f[0], f[1], k, g.nx);
    const float m = 0.5f * (g.mu[k] + g.mu[k + g.nx]);
    f[8][k] = pml(f[8][k], g.inv_dt, g.dx[j], m * dvz_dx);
    f[9][k] = pml(f[9][k], g.inv_dt, dzhv, m * dvx_dz);
}

__device__ __forceinline__ void split_velocity(const Dev2D& g, int j, int k, float dzv, float dzhv) {
    float* const* f = g.f;
    const float r = g.rho[k];
    const float sxx_x = bwd(g, f[4], f[5], k, 1), sxz_z = bwd(g, f[8], f[9], k, g.nx);
    f[0][k] = pml(f[0][k], g.inv_dt, g.dx[j], sxx_x / r);
    f[1][k] = pml(f[1][k], g.inv_dt, dzv, sxz_z / r);
    const float sxz_x = fwd(g, f[8], f[9], k, 1), szz_z = fwd(g, f[6], f[7], k, g.nx);
    const float ra = 0.25f * (g.rho[k] + g.rho[k + 1] + g.rho[k + g.nx] + g.rho[k + g.nx + 1]);
    f[2][k] = pml(f[2][k], g.inv_dt, g.dxh[j], sxz_x / ra);
    f[3][k] = pml(f[3][k], g.inv_dt, dzhv, szz_z / ra);
}

__global__ __launch_bounds__(256) void stress_type1(Dev2D g) {
    count_step(g);
    int i, j;
    if (!cell(g, i, j, kR)) return;
    split_stress(g, j, i * g.nx + j, g.dz[i], g.dzh[i]);
}

__global__ __launch_bounds__(256) void velocity_type1(Dev2D g) {
    int i, j;
    if (!cell(g, i, j, kR)) return;
    split_velocity(g, j, i * g.nx + j, g.dz[i], g.dzh[i]);
}

// updateType2StressKernel (:315-360): plain update inside, split PML in the frame.
__global__ __launch_bounds__(256) void stress_type2(Dev2D g) {
    count_step(g);
    int i, j;
    if (!cell(g, i, j, kR)) return;
    const int k = i * g.nx + j;
    float* const* f = g.f;
    const float dvz_dx = bwd1(g, f[11], k, 1), dvx_dz = fwd1(g, f[10], k, g.nx);
    const float dvx_dx = fwd1(g, f[10], k, 1), dvz_dz = bwd1(g, f[11], k, g.nx);
    const bool frame = i < g.fs || i >= g.nz - g.fs || j < g.fs || j >= g.nx - g.fs;
    if (!frame) {
        f[14][k] += g.dt * g.mu_z[k] * (dvz_dx + dvx_dz);
        f[12][k] += g.dt * (g.l2m_x[k] * dvx_dx + g.lam_x[k] * dvz_dz);
        f[13][k] += g.dt * (g.lam_x[k] * dvx_dx + g.l2m_x[k] * dvz_dz);
    } else {
        const float dz = g.dz[i], dzh = g.dzh[i], dx = g.dx[j], dxh = g.dxh[j];
        f[8][k] = pml(f[8][k], g.inv_dt, dx, g.mu_z[k] * dvz_dx);
        f[9][k] = pml(f[9][k], g.inv_dt, dzh, g.mu_z[k] * dvx_dz);
        f[4][k] = pml(f[4][k], g.inv_dt, dxh, g.l2m_x[k] * dvx_dx);
        f[6][k] = pml(f[6][k], g.inv_dt, dxh, g.lam_x[k] * dvx_dx);
        f[5][k] = pml(f[5][k], g.inv_dt, dz, g.lam_x[k] * dvz_dz);
        f[7][k] = pml(f[7][k], g.inv_dt, dz, g.l2m_x[k] * dvz_dz);
        f[14][k] = f[8][k] + f[9][k];
        f[12][k] = f[4][k] + f[5][k];
        f[13][k] = f[6][k] + f[7][k];
    }
}

// updateType2VelocityKernel (:364-412).
__device__ __forceinline__ void hybrid_velocity(const Dev2D& g, int i, int j) {
    const int k = i * g.nx + j;
    float* const* f = g.f;
    const float sxz_x = fwd1(g, f[14], k, 1), szz_z = fwd1(g, f[13], k, g.nx);
    const float sxx_x = bwd1(g, f[12], k, 1), sxz_z = bwd1(g, f[14], k, g.nx);
    const bool frame = i < g.fs || i >= g.nz - g.fs || j < g.fs || j >= g.nx - g.fs;
    if (!frame) {
        f[11][k] += g.dt * (sxz_x + szz_z) / g.rho_xz[k];
        f[10][k] += g.dt * (sxx_x + sxz_z) / g.rho_o[k];
    } else {
        const float dz = g.dz[i], dzh = g.dzh[i], dx = g.dx[j], dxh = g.dxh[j];
        f[2][k] = pml(f[2][k], g.inv_dt, dxh, sxz_x / g.rho_xz[k]);
        f[3][k] = pml(f[3][k], g.inv_dt, dzh, szz_z / g.rho_xz[k]);
        f[0][k] = pml(f[0][k], g.inv_dt, dx, sxx_x / g.rho_o[k]);
        f[1][k] = pml(f[1][k], g.inv_dt, dz, sxz_z / g.rho_o[k]);
        f[10][k] = f[0][k] + f[1][k];
        f[11][k] = f[2][k] + f[3][k];
    }
}

__global__ __launch_bounds__(256) void velocity_type2(Dev2D g) {
    int i, j;
    if (cell(g, i, j, kR)) hybrid_velocity(g, i, j);
}

// updateType3StressKernel (:416-487), free surface on.
__device__ __forceinline__ void surface_stress(const Dev2D& g, int i, int j) {
    const int k = i * g.nx + j;
    const bool near = i <= g.fs + 5 && g.up;
    const float dzv = near ? 0.0f : g.dz[i], dzhv = near ? 0.0f : g.dzh[i];
    if (i != g.fs) {
        split_stress(g, j, k, dzv, dzhv);
        return;
    }
    float* const* f = g.f;
    f[6][k] = 0.0f;
    f[7][k] = 0.0f;
    const float dvx_dx = fwd(g, f[0], f[1], k, 1);
    f[4][k] = pml(f[4][k], g.inv_dt, g.dxh[j], g.dp[j] * dvx_dx);
    f[5][k] = 0.0f;
    const float dvz_dx = g.c1 * sdiff(f[2], f[3], k, k - 1) + g.c2 * sdiff(f[2], f[3], k + 1, k - 2);
    const float dvx_dz = g.c1 * sdiff(f[0], f[1], k + g.nx, k) + g.c2 * sdiff(f[0], f[1], k + 2 * g.nx, k - g.nx);
    const float m = 0.5f * (g.mu[k] + g.mu[k + g.nx]);
    f[8][k] = pml(f[8][k], g.inv_dt, g.dx[j], m * dvz_dx);
    f[9][k] = pml(f[9][k], g.inv_dt, dzhv, m * dvx_dz);
}

// mirrorType3StressKernel (:491-508) for column j.
__device__ __forceinline__ void mirror_column(const Dev2D& g, int j) {
    float* const* f = g.f;
    for (int l = 1; l <= 4; ++l) {
        const int above = (g.fs - l) * g.nx + j, below = (g.fs + l) * g.nx + j;
        if (above >= 0 && below < g.n) {
            f[6][above] = -f[6][below];
            f[7][above] = -f[7][below];
            const int shear = (g.fs + l - 1) * g.nx + j;
            f[8][above] = -f[8][shear];
            f[9][above] = -f[9][shear];
        }
    }
}

__global__ __launch_bounds__(256) void stress_type3(Dev2D g) {
    count_step(g);
    int i, j;
    if (cell(g, i, j, g.fs)) surface_stress(g, i, j);
}

__global__ void mirror_type3(Dev2D g) {
    const int j = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= kR && j < g.nx - kR) mirror_column(g, j);
}

// updateType3VelocityKernel (:512-563): air rows above the surface are zeroed.
__device__ __forceinline__ void surface_velocity(const Dev2D& g, int i, int j) {
    const int k = i * g.nx + j;
    if (i < g.fs) {
        g.f[0][k] = 0.0f; g.f[1][k] = 0.0f; g.f[2][k] = 0.0f; g.f[3][k] = 0.0f;
        return;
    }
    const bool near = i <= g.fs + 5 && g.up;
    split_velocity(g, j, k, near ? 0.0f : g.dz[i], near ? 0.0f : g.dzh[i]);
}

__global__ __launch_bounds__(256) void velocity_type3(Dev2D g) {
    int i, j;
    if (cell(g, i, j, kR)) surface_velocity(g, i, j);
}

// applyPointSource (:567-636); `atomic` for overlapping multi-source injection.
template <bool atomic>
__device__ void point_source(const Dev2D& g, float value, int idx, int kind, float sin_a, float cos_a) {
    if (idx < 0 || idx >= g.n || !isfinite(value)) return;
    const int sx = idx % g.nx, sz = idx / g.nx;
    if (sx < kR || sx >= g.nx - kR || sz < kR || sz >= g.nz - kR || (g.type == 3 && g.up && sz <= g.fs)) return;
    const bool split = g.type == 1 || g.type == 3;
    auto add = [](float* p, float v) {
        if (atomic) atomicAdd(p, v);
        else *p += v;
    };
    auto rho_avg = [&](int k) { return 0.25f * (g.rho[k] + g.rho[k + 1] + g.rho[k + g.nx] + g.rho[k + g.nx + 1]); };
    if (kind == 0) {
        if (split) add(g.f[2] + idx, value * g.dt / rho_avg(idx));
        else if (g.type == 2) add(g.f[11] + idx, value * g.dt / g.rho_xz[idx]);
    } else if (kind == 1) {
        if (split) add(g.f[0] + idx, value * g.dt / g.rho[idx]);
        else if (g.type == 2) add(g.f[10] + idx, value * g.dt / g.rho_o[idx]);
    } else if (kind == 2) {
        if (split) {
            add(g.f[4] + idx, value * 0.5f); add(g.f[5] + idx, value * 0.5f);
            add(g.f[6] + idx, value * 0.5f); add(g.f[7] + idx, value * 0.5f);
        } else if (g.type == 2) {
            add(g.f[12] + idx, value); add(g.f[13] + idx, value);
        }
    } else if (kind == 3) {
        if (split) { add(g.f[8] + idx, value * 0.5f); add(g.f[9] + idx, value * 0.5f); }
        else if (g.type == 2) add(g.f[14] + idx, value);
    } else if (kind == 4) {
        const float fx = sin_a * value, fz = cos_a * value;
        if (split) {
            add(g.f[0] + idx, fx * g.dt / g.rho[idx]);
            add(g.f[2] + idx, fz * g.dt / rho_avg(idx));
        } else if (g.type == 2) {
            add(g.f[10] + idx, fx * g.dt / g.rho_o[idx]);
            add(g.f[11] + idx, fz * g.dt / g.rho_xz[idx]);
        }
    }
}

// extractReceiverSamplesKernel (:793-818) for receiver r; trace mode writes
// column *counter - first.
__device__ __forceinline__ void sample_receiver(const Dev2D& g, const Rec2D& R, int r) {
    const int k = R.idx[r];
    const long long o = R.len ? static_cast<long long>(r) * R.len + (*g.counter - R.first) : r;
    if (k < 0 || k >= g.n) {
        R.vx[o] = 0.0f;
        R.vz[o] = 0.0f;
        return;
    }
    const bool split = g.type == 1 || g.type == 3;
    R.vx[o] = split ? g.f[0][k] + g.f[1][k] : g.f[10][k];
    R.vz[o] = split ? g.f[2][k] + g.f[3][k] : g.f[11][k];
}

// One workgroup after the velocity half-step, in the reference's order: the
// source(s) (injectSourceKernel :638-640 or injectMultipleSourcesKernel
// :822-834 with host-evaluated Ricker values), zeroBoundaryKernel (:643-667),
// then the receiver samples.  Sources are interior cells and the zeroed frame
// is not, so the two touch disjoint cells; the receivers read after both.
constexpr int kTailThreads = 256;

__global__ __launch_bounds__(kTailThreads) void step_tail(Dev2D g, Tail2D T) {
    const int tid = int(threadIdx.x), nt = kTailThreads;
    if (T.count > 0) {
        for (int r = tid; r < T.count; r += nt)
            point_source<true>(g, T.list[r].value, T.list[r].idx, T.kind, T.sin_a, T.cos_a);
    } else if (tid == 0 && T.single.active) {
        const float v = T.single.table ? T.single.table[*g.counter] : T.single.value;
        point_source<false>(g, v, T.single.idx, T.single.kind, T.single.sin_a, T.single.cos_a);
    }
    const bool split = g.type == 1 || g.type == 3;
    auto zero = [&](int k) {
        if (split) { g.f[0][k] = 0.0f; g.f[1][k] = 0.0f; g.f[2][k] = 0.0f; g.f[3][k] = 0.0f; }
        else { g.f[10][k] = 0.0f; g.f[11][k] = 0.0f; }
    };
    for (int c = tid; c < g.nx; c += nt) { zero(c); zero((g.nz - 1) * g.nx + c); }
    for (int r = tid; r < g.nz; r += nt) { zero(r * g.nx); zero(r * g.nx + (g.nx - 1)); }
    __syncthreads();
    for (int r = tid; r < T.rec.count; r += nt) sample_receiver(g, T.rec, r);
}

// Receiver samples on their own (the per-step extractReceiverSamples call).
__global__ void receivers(Dev2D g, Rec2D R) {
    const int r = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (r < R.count) sample_receiver(g, R, r);
}

dim3 grid_of(const Dev2D& g) { return dim3(unsigned((g.nx + 63) / 64), unsigned((g.nz + 3) / 4)); }

}  // namespace

// runGpuSimulationStep's scheme choice (:1036-1062): type 3 with the surface off runs type 1.
hipError_t launch_stress2d(const Dev2D& g, hipStream_t s) {
    if (g.type == 2) hipLaunchKernelGGL(stress_type2, grid_of(g), dim3(64, 4), 0, s, g);
    else if (g.type == 3 && g.up) {
        hipLaunchKernelGGL(stress_type3, grid_of(g), dim3(64, 4), 0, s, g);
        hipLaunchKernelGGL(mirror_type3, dim3(unsigned((g.nx + 255) / 256)), dim3(256), 0, s, g);
    } else hipLaunchKernelGGL(stress_type1, grid_of(g), dim3(64, 4), 0, s, g);
    return hipGetLastError();
}

hipError_t launch_velocity2d(const Dev2D& g, const Tail2D& tail, hipStream_t s) {
    if (g.type == 2) hipLaunchKernelGGL(velocity_type2, grid_of(g), dim3(64, 4), 0, s, g);
    else if (g.type == 3 && g.up) hipLaunchKernelGGL(velocity_type3, grid_of(g), dim3(64, 4), 0, s, g);
    else hipLaunchKernelGGL(velocity_type1, grid_of(g), dim3(64, 4), 0, s, g);
    hipLaunchKernelGGL(step_tail, dim3(1), dim3(kTailThreads), 0, s, g, tail);
    return hipGetLastError();
}

hipError_t launch_receivers2d(const Dev2D& g, const Rec2D& r, hipStream_t s) {
    if (r.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(receivers, dim3(unsigned((r.count + 127) / 128)), dim3(128), 0, s, g, r);
    return hipGetLastError();
}

}  // namespace seis2d
