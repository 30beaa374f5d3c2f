// seis3d_kernels.hip - CDNA4 (gfx950) kernels of the 3D O(8) velocity-stress
// step: fused velocity and stress half-steps (CPML psi memories and source
// injection fused in), free-surface ghost planes, reflecting walls,
// receivers, materials and snapshots.
//
// The direct kernel family (one thread per updated cell, global loads only: the
// reference's decomposition of velocityKernel3D / stressKernel3D,
// cuda/Cuda_Seis3D.cu:222-310) lives here; it computes the same arithmetic in the
// same order as the LDS-DMA 2.5D family of seis3d_dma.hip, the default.
// Built with -ffp-contract=off: every multiply and add rounds separately,
// as in the FP32 CPU oracle.
#include "seis3d_device.hpp"
#include "seis3d_kernels.hpp"

namespace seis3d {

// ---------------------------------------------------------------------------
// Helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ long long cell(const DevGrid& g, int x, int y, int z) {
    return (static_cast<long long>(z) * g.gy + y) * g.px + x;
}

template <bool plus>
__device__ __forceinline__ float deriv_global(const float* __restrict__ f, long long i, long long s, float invh) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = f[i + (j - (plus ? 3 : 4)) * s];
    return stencil8(v, invh);
}

// Source weight index r^2 for a cell inside the 3x3x3 box, -1 outside.
__device__ __forceinline__ int source_r2(const DevSource& s, int x, int y, int z) {
    const int dx = x - s.x, dy = y - s.y, dz = z - s.z;
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return -1;
    return dx * dx + dy * dy + dz * dz;
}

// ---------------------------------------------------------------------------
// Direct family
// ---------------------------------------------------------------------------

// velocityKernel3D (cuda/Cuda_Seis3D.cu:222-245) with the force source of
// sourceKernel3D (:326-330) fused as an epilogue.
__global__ __launch_bounds__(256) void velocity_direct(DevStep P) {
    const DevGrid& g = P.g;
    const int x = int(blockIdx.x * blockDim.x + threadIdx.x) + 4;
    const int y = int(blockIdx.y * blockDim.y + threadIdx.y) + 4;
    const int z = int(blockIdx.z) + 4;
    if (x >= g.gx - 4 || y >= g.gy - 4 || z >= g.gz - 4) return;
    const long long i = cell(g, x, y, z);
    const long long st[3] = {1, g.px, g.plane};
    const int qx = slab_q(x, g.gx, g.pml, false);
    const int qy = slab_q(y, g.gy, g.pml, false);
    const int qz = slab_q(z, g.gz, g.pml, g.free_surface != 0);
    const int sf[3][3] = {{3, 6, 8}, {6, 4, 7}, {8, 7, 5}};
    const int r2 = P.src.active ? source_r2(P.src, x, y, z) : -1;
    bool bad = false;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        float div = 0.0f;
#pragma unroll
        for (int axis = 0; axis < 3; ++axis) {
            const float* fld = P.p.f[sf[v][axis]];
            float d = axis == v ? deriv_global<true>(fld, i, st[axis], P.invh)
                                : deriv_global<false>(fld, i, st[axis], P.invh);
            const int half = axis == v ? 1 : 0;
            if (axis == 0 && qx >= 0)
                d = cpml_update(P.p.psi[0], psi_index_x(g, v, qx, y, z), P.p.coef[0] + (half * g.gx + x) * 3, d);
            if (axis == 1 && qy >= 0)
                d = cpml_update(P.p.psi[1], psi_index_y(g, v, qy, x, z), P.p.coef[1] + (half * g.gy + y) * 3, d);
            if (axis == 2 && qz >= 0)
                d = cpml_update(P.p.psi[2], psi_index_z(g, v, qz, x, y), P.p.coef[2] + (half * g.gz + z) * 3, d);
            div += d;
        }
        const float h = face_buoyancy(P.p.inv_rho[i], P.p.inv_rho[i + st[v]]);
        float value = P.p.f[v][i] + P.dt * h * div;
        if (r2 >= 0) value += (P.src.scale * P.src.w[r2] / P.src.norm) * P.src.dir[v] * h;
        P.p.f[v][i] = value;
        bad |= !isfinite(value);
    }
    if (bad) atomicOr(P.p.flags, 1u);
}

// stressKernel3D (cuda/Cuda_Seis3D.cu:281-310) with the explosion source of
// sourceKernel3D (:323-324) fused as an epilogue.
__global__ __launch_bounds__(256) void stress_direct(DevStep P) {
    const DevGrid& g = P.g;
    const int x = int(blockIdx.x * blockDim.x + threadIdx.x) + 4;
    const int y = int(blockIdx.y * blockDim.y + threadIdx.y) + 4;
    const int z = int(blockIdx.z) + 4;
    if (x >= g.gx - 4 || y >= g.gy - 4 || z >= g.gz - 4) return;
    const long long i = cell(g, x, y, z);
    const long long st[3] = {1, g.px, g.plane};
    const int qx = slab_q(x, g.gx, g.pml, false);
    const int qy = slab_q(y, g.gy, g.pml, false);
    const int qz = slab_q(z, g.gz, g.pml, g.free_surface != 0);
    float d[3][3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float raw = a == v ? deriv_global<false>(P.p.f[v], i, st[a], P.invh)
                               : deriv_global<true>(P.p.f[v], i, st[a], P.invh);
            const int half = a != v ? 1 : 0;
            if (a == 0 && qx >= 0)
                raw = cpml_update(P.p.psi[0], psi_index_x(g, 3 + v, qx, y, z), P.p.coef[0] + (half * g.gx + x) * 3, raw);
            if (a == 1 && qy >= 0)
                raw = cpml_update(P.p.psi[1], psi_index_y(g, 3 + v, qy, x, z), P.p.coef[1] + (half * g.gy + y) * 3, raw);
            if (a == 2 && qz >= 0)
                raw = cpml_update(P.p.psi[2], psi_index_z(g, 3 + v, qz, x, y), P.p.coef[2] + (half * g.gz + z) * 3, raw);
            d[v][a] = raw;
        }
    }
    const float lam = P.p.lambda[i], mu = P.p.mu[i];
    const bool surface = g.free_surface && z == 4;
    if (surface) d[2][2] = -lam / (lam + 2 * mu) * (d[0][0] + d[1][1]);
    const float trace = d[0][0] + d[1][1] + d[2][2];
    const float* r = P.p.inv_mu;
    float sxx = P.p.f[3][i] + P.dt * (lam * trace + 2 * mu * d[0][0]);
    float syy = P.p.f[4][i] + P.dt * (lam * trace + 2 * mu * d[1][1]);
    float szz = surface ? 0.0f : P.p.f[5][i] + P.dt * (lam * trace + 2 * mu * d[2][2]);
    const float sxy = P.p.f[6][i] + P.dt * shear_from_inverse(r[i], r[i + st[0]], r[i + st[1]], r[i + st[0] + st[1]]) * (d[0][1] + d[1][0]);
    const float syz = P.p.f[7][i] + P.dt * shear_from_inverse(r[i], r[i + st[1]], r[i + st[2]], r[i + st[1] + st[2]]) * (d[1][2] + d[2][1]);
    const float sxz = P.p.f[8][i] + P.dt * shear_from_inverse(r[i], r[i + st[0]], r[i + st[2]], r[i + st[0] + st[2]]) * (d[0][2] + d[2][0]);
    if (P.src.active) {
        const int r2 = source_r2(P.src, x, y, z);
        if (r2 >= 0) {
            const float value = P.src.scale * P.src.w[r2] / P.src.norm;
            sxx += value; syy += value; szz += value;
        }
    }
    P.p.f[3][i] = sxx; P.p.f[4][i] = syy; P.p.f[5][i] = szz;
    P.p.f[6][i] = sxy; P.p.f[7][i] = syz; P.p.f[8][i] = sxz;
    const bool bad = !isfinite(sxx) || !isfinite(syy) || !isfinite(szz) || !isfinite(sxy) || !isfinite(syz) || !isfinite(sxz);
    if (bad) atomicOr(P.p.flags, 1u);
}

// ---------------------------------------------------------------------------
// Boundary planes
// ---------------------------------------------------------------------------

// stressGhostKernel3D (cuda/Cuda_Seis3D.cu:203-219).
__global__ void stress_ghost(DevStep P) {
    const DevGrid& g = P.g;
    const int x = int(blockIdx.x * blockDim.x + threadIdx.x) + 4;
    const int y = int(blockIdx.y * blockDim.y + threadIdx.y) + 4;
    if (x >= g.gx - 4 || y >= g.gy - 4) return;
    float* const* f = P.p.f;
    f[5][cell(g, x, y, 4)] = 0.0f;
#pragma unroll
    for (int depth = 1; depth <= 4; ++depth) {
        const long long ghost = cell(g, x, y, 4 - depth), mirror = cell(g, x, y, 4 + depth);
        const long long shear_mirror = cell(g, x, y, 3 + depth);
        f[5][ghost] = -f[5][mirror];
        f[8][ghost] = -f[8][shear_mirror];
        f[7][ghost] = -f[7][shear_mirror];
        f[3][ghost] = f[3][mirror];
        f[4][ghost] = f[4][mirror];
        f[6][ghost] = f[6][mirror];
    }
}

// velocityGhostKernel3D (cuda/Cuda_Seis3D.cu:249-272).
__device__ __forceinline__ void velocity_ghost_at(const DevStep& P, float spacing, int x, int y) {
    const DevGrid& g = P.g;
    float* const* f = P.p.f;
    const long long i = cell(g, x, y, 4);
    const float invh = P.invh;
    const float ex = deriv_global<false>(f[0], i, 1, invh);
    const float ey = deriv_global<false>(f[1], i, g.px, invh);
    const long long next = i + g.plane;
    const float dz_vx = -(1.5f * deriv_global<true>(f[2], i, 1, invh) - 0.5f * deriv_global<true>(f[2], next, 1, invh));
    const float dz_vy = -(1.5f * deriv_global<true>(f[2], i, g.px, invh) - 0.5f * deriv_global<true>(f[2], next, g.px, invh));
    const float lam = P.p.lambda[i], mu = P.p.mu[i];
    const float dz_vz = -lam / (lam + 2 * mu) * (ex + ey);
#pragma unroll
    for (int depth = 1; depth <= 4; ++depth) {
        const long long ghost = cell(g, x, y, 4 - depth), mirror = cell(g, x, y, 4 + depth);
        f[0][ghost] = f[0][mirror] - float(2 * depth) * spacing * dz_vx;
        f[1][ghost] = f[1][mirror] - float(2 * depth) * spacing * dz_vy;
        f[2][ghost] = f[2][cell(g, x, y, 3 + depth)] - float(2 * depth - 1) * spacing * dz_vz;
    }
}

__global__ void velocity_ghost(DevStep P, float spacing) {
    const int x = int(blockIdx.x * blockDim.x + threadIdx.x) + 4;
    const int y = int(blockIdx.y * blockDim.y + threadIdx.y) + 4;
    if (x >= P.g.gx - 4 || y >= P.g.gy - 4) return;
    velocity_ghost_at(P, spacing, x, y);
}

// The velocity ghosts of several shots in one launch (blockIdx.z = shot; shots without a free
// surface return at once).
__global__ void velocity_ghost_shots(const DevStep* __restrict__ steps, const float* __restrict__ spacing) {
    const int s = int(blockIdx.z);
    const DevStep& P = steps[s];
    if (!P.g.free_surface) return;
    const int x = int(blockIdx.x * blockDim.x + threadIdx.x) + 4;
    const int y = int(blockIdx.y * blockDim.y + threadIdx.y) + 4;
    if (x >= P.g.gx - 4 || y >= P.g.gy - 4) return;
    velocity_ghost_at(P, spacing[s], x, y);
}

// reflectingGhostKernel3D (cuda/Cuda_Seis3D.cu:152-175), one axis per launch.
__global__ void reflecting_walls(DevStep P, int axis, int velocities) {
    const DevGrid& g = P.g;
    const int lengths[3] = {g.gx, g.gy, g.gz};
    const int u_axis = (axis + 1) % 3, v_axis = (axis + 2) % 3;
    const int u = int(blockIdx.x * blockDim.x + threadIdx.x);
    const int v = int(blockIdx.y * blockDim.y + threadIdx.y);
    if (u >= lengths[u_axis] || v >= lengths[v_axis]) return;
    const int high = lengths[axis] - 4 - 1;
    int p[3] = {0, 0, 0};
    p[u_axis] = u; p[v_axis] = v;
    for (int field = velocities ? 0 : 3; field < (velocities ? 3 : 9); ++field) {
        const bool half = field < 3 ? field == axis : field == 6 ? axis != 2 : field == 7 ? axis != 0
                        : field == 8 ? axis != 1 : false;
        float* f = P.p.f[field];
        for (int side = 0; side < 2; ++side) for (int depth = 1; depth <= 4; ++depth) {
            p[axis] = side ? high + depth : 4 - depth;
            const long long destination = cell(g, p[0], p[1], p[2]);
            if (half && !side && depth == 1) { f[destination] = 0.0f; continue; }
            p[axis] = side ? high + 1 - depth - int(half) : 4 + depth - 1 - int(half);
            const float value = f[cell(g, p[0], p[1], p[2])];
            f[destination] = half ? -value : value;
        }
        if (half) { p[axis] = high; f[cell(g, p[0], p[1], p[2])] = 0.0f; }
    }
}

// ---------------------------------------------------------------------------
// Receivers, materials, snapshots
// ---------------------------------------------------------------------------

// sampleScalar3D (cuda/Cuda_Seis3D.cu:335-352) at a padded index.
__device__ float sample_scalar(const DevStep& P, long long i, int component) {
    const DevGrid& g = P.g;
    float* const* f = P.p.f;
    if (component == 5) return sqrtf((P.p.lambda[i] + 2 * P.p.mu[i]) * P.p.inv_rho[i]);
    if (component == 6) return sqrtf(P.p.mu[i] * P.p.inv_rho[i]);
    if (component == 7) return 1 / P.p.inv_rho[i];
    const float vx = 0.5f * (f[0][i] + f[0][i - 1]);
    const float vy = 0.5f * (f[1][i] + f[1][i - g.px]);
    const float vz = 0.5f * (f[2][i] + f[2][i - g.plane]);
    if (component == 0) return vz;
    if (component == 1) return vx;
    if (component == 2) return vy;
    if (component == 3) return sqrtf(vx * vx + vy * vy + vz * vz);
    if (component == 4) return -(f[3][i] + f[4][i] + f[5][i]) / 3;
    return 0.0f;
}

// receiverKernel3D (cuda/Cuda_Seis3D.cu:421-428): out[3r..3r+2] = vx, vy, vz.
// Receivers arrive as padded coordinates.
__device__ __forceinline__ void receiver_at(const DevStep& P, const int* __restrict__ rec, int r,
                                            float* __restrict__ out) {
    const long long i = cell(P.g, rec[3 * r], rec[3 * r + 1], rec[3 * r + 2]);
    out[3 * r] = sample_scalar(P, i, 1);
    out[3 * r + 1] = sample_scalar(P, i, 2);
    out[3 * r + 2] = sample_scalar(P, i, 0);
}

__global__ void receivers(DevStep P, const int* __restrict__ rec, int count, float* __restrict__ out) {
    const int r = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= count) return;
    receiver_at(P, rec, r, out);
}

// One sample of several shots' receivers in one launch (blockIdx.y = entry).
__global__ void receivers_shots(const DevStep* __restrict__ steps, const ShotSample* __restrict__ samples) {
    const ShotSample& e = samples[blockIdx.y];
    const int r = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= e.count) return;
    receiver_at(steps[e.shot], e.rec, r, e.out);
}

// generateMaterialsKernel3D (cuda/Cuda_Seis3D.cu:114-147).
__global__ void generate_materials(DevStep P, MaterialParams mp) {
    const DevGrid& g = P.g;
    const long long cells = static_cast<long long>(g.gx) * g.gy * g.gz;
    const long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= cells) return;
    const int gxv = int(k % g.gx), gyv = int((k / g.gx) % g.gy), gzv = int(k / (static_cast<long long>(g.gx) * g.gy));
    const int x = max(0, min(mp.nx - 1, gxv - mp.ox));
    const int y = max(0, min(mp.ny - 1, gyv - mp.oy));
    const int z = max(0, min(mp.nz - 1, gzv - mp.oz));
    const float dx = (x - mp.nx * 0.5f) / mp.nx;
    const float dy = (y - mp.ny * 0.5f) / mp.ny;
    const float dz = (z - mp.nz * 0.55f) / mp.nz;
    const float depth = float(z) / float(mp.nz - 1), horizontal = float(x) / float(mp.nx - 1) - 0.5f;
    float blend = 0.0f;
    switch (mp.kind) {
    case 1: blend = depth >= mp.interface_depth; break;
    case 2: blend = dx * dx + dy * dy + dz * dz < mp.anomaly_radius * mp.anomaly_radius; break;
    case 4: blend = depth < 1.0f / 3 ? 0 : depth < 2.0f / 3 ? 0.5f : 1; break;
    case 5: blend = depth >= mp.interface_depth + mp.layer_dip * horizontal; break;
    case 6: blend = fabsf(depth - mp.interface_depth) <= 0.5f * mp.layer_thickness; break;
    case 7: blend = depth >= mp.interface_depth + mp.layer_dip * horizontal + (horizontal > 0 ? mp.fault_throw : 0); break;
    case 8: blend = depth; break;
    case 9: {
        const float vertical = (depth - 0.78f) / 0.52f;
        blend = (dx * dx + dy * dy) / (mp.anomaly_radius * mp.anomaly_radius) + vertical * vertical < 1;
        break;
    }
    default: break;
    }
    const float vp = mp.vp + blend * (mp.layer_vp - mp.vp), vs = mp.vs + blend * (mp.layer_vs - mp.vs);
    const float rho = mp.rho + blend * (mp.layer_rho - mp.rho);
    const long long i = cell(g, gxv, gyv, gzv);
    P.p.inv_rho[i] = 1.0f / rho;
    P.p.mu[i] = rho * vs * vs;
    P.p.lambda[i] = rho * (vp * vp - 2.0f * vs * vs);
}

// extendMaterialsKernel3D (cuda/Cuda_Seis3D.cu:188-200).
__global__ void extend_materials(DevStep P, MaterialParams mp) {
    const DevGrid& g = P.g;
    const long long cells = static_cast<long long>(g.gx) * g.gy * g.gz;
    const long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= cells) return;
    const int x = int(k % g.gx), y = int((k / g.gx) % g.gy), z = int(k / (static_cast<long long>(g.gx) * g.gy));
    const int px = max(mp.ox, min(x, mp.ox + mp.nx - 1));
    const int py = max(mp.oy, min(y, mp.oy + mp.ny - 1));
    const int pz = max(mp.oz, min(z, mp.oz + mp.nz - 1));
    if (x == px && y == py && z == pz) return;
    const long long i = cell(g, x, y, z), j = cell(g, px, py, pz);
    P.p.inv_rho[i] = P.p.inv_rho[j];
    P.p.mu[i] = P.p.mu[j];
    P.p.lambda[i] = P.p.lambda[j];
}

// 1/mu over the padded grid: the reciprocals of shearMu3D
// (cuda/Cuda_Seis3D.cu:275-278), computed once per model instead of four
// times per cell and step.
__global__ void reciprocal_mu(DevStep P) {
    const DevGrid& g = P.g;
    const long long cells = static_cast<long long>(g.gx) * g.gy * g.gz;
    const long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= cells) return;
    const int x = int(k % g.gx), y = int((k / g.gx) % g.gy), z = int(k / (static_cast<long long>(g.gx) * g.gy));
    const long long i = cell(g, x, y, z);
    P.p.inv_mu[i] = 1 / P.p.mu[i];
}

// Packed slice of one collocated component over the display domain
// (extractKernel3D slice part, cuda/Cuda_Seis3D.cu:370-394, without GL).
__global__ void extract_slice(DevStep P, SliceParams sp, float* __restrict__ out) {
    const long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= static_cast<long long>(sp.width) * sp.height) return;
    const int u = int(k % sp.width), v = int(k / sp.width);
    int p[3];
    p[sp.fixed_axis] = sp.fixed;
    p[sp.u_axis] = u + sp.u_origin;
    p[sp.v_axis] = v + sp.v_origin;
    const long long i = cell(P.g, p[0] + sp.ox, p[1] + sp.oy, p[2] + sp.oz);
    const float value = sample_scalar(P, i, sp.component);
    out[k] = isfinite(value) ? value : 0.0f;
    if (!isfinite(value)) atomicOr(P.p.flags, 1u);
}

// Collocated scalar at a fractional padded position by trilinear weights
// (resampleScalar3D, cuda/Cuda_Seis3D.cu:355-367); corners of zero weight are
// not read.
__device__ float resample_scalar(const DevStep& P, float fx, float fy, float fz, int component, float* magnitude) {
    const int ix = int(floorf(fx)), iy = int(floorf(fy)), iz = int(floorf(fz));
    const float fr[3] = {fx - ix, fy - iy, fz - iz};
    float result = 0.0f;
    if (magnitude) *magnitude = 0.0f;
    for (int dz = 0; dz < 2; ++dz)
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const float w = (dx ? fr[0] : 1 - fr[0]) * (dy ? fr[1] : 1 - fr[1]) * (dz ? fr[2] : 1 - fr[2]);
                if (w > 0) {
                    const long long i = cell(P.g, ix + dx, iy + dy, iz + dz);
                    result += w * sample_scalar(P, i, component);
                    if (magnitude) *magnitude += w * sample_scalar(P, i, 3);
                }
            }
    return result;
}

// extractKernel3D (cuda/Cuda_Seis3D.cu:370-418): sample i of the packed view ->
// out[2i] (wave component), out[2i+1] (material component), with statistics.
__global__ __launch_bounds__(256) void extract_views(DevStep P, ViewParams v, long long count, float* __restrict__ out,
                                                     unsigned* __restrict__ stats) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool valid = i < count;
    float wave_max = 0.0f, vel_max = 0.0f, mat_min = 3.402823466e38f, mat_max = 0.0f;
    bool bad = false;
    if (valid) {
        const int nx = v.domain[0], ny = v.domain[1], nz = v.domain[2];
        int x, y, z;
        float fx = 0, fy = 0, fz = 0;
        const bool in_volume = i >= v.volume_offset;
        if (i < v.slice_offset[1]) {
            x = int(i % nx) + v.origin[0]; y = v.slices[1]; z = int(i / nx) + v.origin[2];
        } else if (i < v.slice_offset[2]) {
            const long long j = i - v.slice_offset[1];
            x = v.slices[0]; y = int(j % ny) + v.origin[1]; z = int(j / ny) + v.origin[2];
        } else if (!in_volume) {
            const long long j = i - v.slice_offset[2];
            x = int(j % nx) + v.origin[0]; y = int(j / nx) + v.origin[1]; z = v.slices[2];
        } else {
            const long long j = i - v.volume_offset;
            fx = (float(j % v.volume[0]) + .5f) * float(nx) / v.volume[0] - .5f + v.origin[0];
            fy = (float((j / v.volume[0]) % v.volume[1]) + .5f) * float(ny) / v.volume[1] - .5f + v.origin[1];
            fz = (float(j / (static_cast<long long>(v.volume[0]) * v.volume[1])) + .5f) * float(nz) / v.volume[2] - .5f +
                 v.origin[2];
            x = int(floorf(fx)); y = int(floorf(fy)); z = int(floorf(fz));
        }
        float value, material, magnitude;
        if (in_volume) {
            value = resample_scalar(P, fx + v.ox, fy + v.oy, fz + v.oz, v.component, &magnitude);
            material = resample_scalar(P, fx + v.ox, fy + v.oy, fz + v.oz, v.material_component, nullptr);
        } else {
            const long long c = cell(P.g, x + v.ox, y + v.oy, z + v.oz);
            value = sample_scalar(P, c, v.component);
            magnitude = sample_scalar(P, c, 3);
            material = sample_scalar(P, c, v.material_component);
        }
        out[2 * i] = isfinite(value) ? value : 0.0f;
        out[2 * i + 1] = isfinite(material) ? material : 0.0f;
        bad = !isfinite(value) || !isfinite(material) || !isfinite(magnitude) || material < 0;
        const bool physical = x >= 0 && x < v.nx && y >= 0 && y < v.ny && z >= 0 && z < v.nz;
        if ((physical || v.include_boundary) && isfinite(value)) wave_max = fabsf(value);
        if ((physical || v.include_boundary) && isfinite(magnitude)) vel_max = magnitude;
        if (physical && isfinite(material)) { mat_min = material; mat_max = material; }
    }
    // Whole-wave reduction (every lane takes part; lanes past the end hold neutral values).
    for (int off = 32; off > 0; off >>= 1) {
        wave_max = fmaxf(wave_max, __shfl_xor(wave_max, off));
        vel_max = fmaxf(vel_max, __shfl_xor(vel_max, off));
        mat_min = fminf(mat_min, __shfl_xor(mat_min, off));
        mat_max = fmaxf(mat_max, __shfl_xor(mat_max, off));
    }
    if (bad) atomicOr(stats + 1, 1u);
    if ((threadIdx.x & 63) == 0) {
        atomicMax(stats + 0, __float_as_uint(wave_max));
        atomicMin(stats + 2, __float_as_uint(mat_min));
        atomicMax(stats + 3, __float_as_uint(mat_max));
        atomicMax(stats + 4, __float_as_uint(vel_max));
    }
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

hipError_t launch_velocity(const DevStep& P, const KernelConfig& k, hipStream_t s) {
    if (k.family == 4) return launch_dma(true, P, k.tile_y, k.z_chunk, s);
    if (out_of_place(P.p)) return hipErrorInvalidValue;   // in-place families only
    const dim3 block(64, 4), grid((P.g.gx - 8 + 63) / 64, (P.g.gy - 8 + 3) / 4, P.g.gz - 8);
    hipLaunchKernelGGL(velocity_direct, grid, block, 0, s, P);
    return hipGetLastError();
}

hipError_t launch_stress(const DevStep& P, const KernelConfig& k, hipStream_t s) {
    if (k.family == 4) return launch_dma(false, P, k.tile_y, k.z_chunk, s);
    if (out_of_place(P.p)) return hipErrorInvalidValue;
    const dim3 block(64, 4), grid((P.g.gx - 8 + 63) / 64, (P.g.gy - 8 + 3) / 4, P.g.gz - 8);
    hipLaunchKernelGGL(stress_direct, grid, block, 0, s, P);
    return hipGetLastError();
}

hipError_t launch_stress_ghost(const DevStep& P, hipStream_t s) {
    const dim3 block(64, 4), grid((P.g.gx - 8 + 63) / 64, (P.g.gy - 8 + 3) / 4);
    hipLaunchKernelGGL(stress_ghost, grid, block, 0, s, P);
    return hipGetLastError();
}

hipError_t launch_velocity_ghost(const DevStep& P, float spacing, hipStream_t s) {
    const dim3 block(64, 4), grid((P.g.gx - 8 + 63) / 64, (P.g.gy - 8 + 3) / 4);
    hipLaunchKernelGGL(velocity_ghost, grid, block, 0, s, P, spacing);
    return hipGetLastError();
}

hipError_t launch_velocity_ghost_shots(const DevStep* steps, const float* spacing, int nshots, int max_gx, int max_gy,
                                       hipStream_t s) {
    if (nshots <= 0) return hipSuccess;
    const dim3 block(64, 4), grid((max_gx - 8 + 63) / 64, (max_gy - 8 + 3) / 4, nshots);
    hipLaunchKernelGGL(velocity_ghost_shots, grid, block, 0, s, steps, spacing);
    return hipGetLastError();
}

hipError_t launch_receivers_shots(const DevStep* steps, const ShotSample* samples, int entries, int max_count,
                                  hipStream_t s) {
    if (entries <= 0 || max_count <= 0) return hipSuccess;
    hipLaunchKernelGGL(receivers_shots, dim3((max_count + 63) / 64, entries), dim3(64), 0, s, steps, samples);
    return hipGetLastError();
}

hipError_t launch_reflecting_walls(const DevStep& P, bool velocities, hipStream_t s) {
    const int lengths[3] = {P.g.gx, P.g.gy, P.g.gz};
    const dim3 block(16, 16);
    for (int axis = 0; axis < 3; ++axis) {
        const dim3 grid((lengths[(axis + 1) % 3] + 15) / 16, (lengths[(axis + 2) % 3] + 15) / 16);
        hipLaunchKernelGGL(reflecting_walls, grid, block, 0, s, P, axis, velocities ? 1 : 0);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_receivers(const DevStep& P, const int* rec, int count, float* out, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(receivers, dim3((count + 63) / 64), dim3(64), 0, s, P, rec, count, out);
    return hipGetLastError();
}

hipError_t launch_generate_materials(const DevStep& P, const MaterialParams& mp, hipStream_t s) {
    const long long cells = static_cast<long long>(P.g.gx) * P.g.gy * P.g.gz;
    hipLaunchKernelGGL(generate_materials, dim3(unsigned((cells + 255) / 256)), dim3(256), 0, s, P, mp);
    return hipGetLastError();
}

hipError_t launch_extend_materials(const DevStep& P, const MaterialParams& mp, hipStream_t s) {
    const long long cells = static_cast<long long>(P.g.gx) * P.g.gy * P.g.gz;
    hipLaunchKernelGGL(extend_materials, dim3(unsigned((cells + 255) / 256)), dim3(256), 0, s, P, mp);
    return hipGetLastError();
}

hipError_t launch_reciprocal_mu(const DevStep& P, hipStream_t s) {
    const long long cells = static_cast<long long>(P.g.gx) * P.g.gy * P.g.gz;
    hipLaunchKernelGGL(reciprocal_mu, dim3(unsigned((cells + 255) / 256)), dim3(256), 0, s, P);
    return hipGetLastError();
}

hipError_t launch_extract_views(const DevStep& P, const ViewParams& vp, long long count, float* out, unsigned* stats,
                                hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(extract_views, dim3(unsigned((count + 255) / 256)), dim3(256), 0, s, P, vp, count, out, stats);
    return hipGetLastError();
}

hipError_t launch_extract_slice(const DevStep& P, const SliceParams& sp, float* out, hipStream_t s) {
    const long long n = static_cast<long long>(sp.width) * sp.height;
    hipLaunchKernelGGL(extract_slice, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, P, sp, out);
    return hipGetLastError();
}

}  // namespace seis3d
