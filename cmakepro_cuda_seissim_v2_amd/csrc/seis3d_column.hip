// seis3d_column.hip - column-marching velocity and stress half-steps.
//
// Same arithmetic, operation for operation, as velocityKernel3D /
// stressKernel3D (cuda/Cuda_Seis3D.cu:222-310).  Each thread owns one (x, y)
// column of a 64 x BY tile and marches z through a chunk of planes: the
// z-derivative fields live in 9-deep register queues (one new plane loaded
// per step instead of eight), while the x/y stencils are read straight from
// global memory, where the tile's own rows and the 4-cell halos are served by
// the vector L1 and the XCD's L2.  No LDS and no workgroup barriers, so
// occupancy and independent waves hide the memory latency.
#include "seis3d_device.hpp"
#include "seis3d_kernels.hpp"

namespace seis3d {

namespace {

__device__ __forceinline__ long long ccell(const DevGrid& g, int x, int y, int z) {
    return (static_cast<long long>(z) * g.gy + y) * g.px + x;
}

__device__ __forceinline__ int source_r2c(const DevSource& s, int x, int y, int z) {
    const int dx = x - s.x, dy = y - s.y, dz = z - s.z;
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return -1;
    return dx * dx + dy * dy + dz * dz;
}

template <bool plus>
__device__ __forceinline__ float dglob(const float* __restrict__ f, long long i, long long s, float invh) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = f[i + (j - (plus ? 3 : 4)) * s];
    return stencil8(v, invh);
}

}  // namespace

template <int BY>
__global__ __launch_bounds__(64 * BY) void velocity_column(DevStep P, TileMap M) {
    int bx, by, bz;
    if (!tile_of(M, bx, by, bz)) return;
    const DevGrid& g = P.g;
    const int x = 4 + bx * 64 + int(threadIdx.x), y = 4 + by * BY + int(threadIdx.y);
    if (x >= g.gx - 4 || y >= g.gy - 4) return;
    int z0, z1;
    chunk_of(M, g.gz, bz, z0, z1);
    const int qx = slab_q(x, g.gx, g.pml, false);
    const int qy = slab_q(y, g.gy, g.pml, false);
    const float invh = P.invh;
    const float* sxx = P.p.f[3];
    const float* syy = P.p.f[4];
    const float* szz = P.p.f[5];
    const float* sxy = P.p.f[6];
    const float* syz = P.p.f[7];
    const float* sxz = P.p.f[8];
    const long long col = ccell(g, x, y, 0);

    float q_xz[9], q_yz[9], q_zz[9];   // columns z-4..z+4
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const long long k = col + static_cast<long long>(z0 - 4 + j) * g.plane;
        q_xz[j + 1] = sxz[k];
        q_yz[j + 1] = syz[k];
        q_zz[j + 1] = szz[k];
    }
    bool bad = false;
    for (int z = z0; z < z1; ++z) {
        const long long i = col + static_cast<long long>(z) * g.plane;
        const long long lead = i + 4 * g.plane;
        shift(q_xz, sxz[lead]);
        shift(q_yz, syz[lead]);
        shift(q_zz, szz[lead]);
        const int qz = slab_q(z, g.gz, g.pml, g.free_surface != 0);
        float w[8];
        // vx: D+x sxx (half), D-y sxy, D-z sxz.
        float d = dglob<true>(sxx, i, 1, invh);
        if (qx >= 0) d = cpml_update(P.p.psi[0], psi_index_x(g, 0, qx, y, z), P.p.coef[0] + (g.gx + x) * 3, d);
        float div_x = 0.0f;
        div_x += d;
        d = dglob<false>(sxy, i, g.px, invh);
        if (qy >= 0) d = cpml_update(P.p.psi[1], psi_index_y(g, 0, qy, x, z), P.p.coef[1] + y * 3, d);
        div_x += d;
#pragma unroll
        for (int j = 0; j < 8; ++j) w[j] = q_xz[j];
        d = stencil8(w, invh);
        if (qz >= 0) d = cpml_update(P.p.psi[2], psi_index_z(g, 0, qz, x, y), P.p.coef[2] + z * 3, d);
        div_x += d;
        // vy: D-x sxy, D+y syy (half), D-z syz.
        d = dglob<false>(sxy, i, 1, invh);
        if (qx >= 0) d = cpml_update(P.p.psi[0], psi_index_x(g, 1, qx, y, z), P.p.coef[0] + x * 3, d);
        float div_y = 0.0f;
        div_y += d;
        d = dglob<true>(syy, i, g.px, invh);
        if (qy >= 0) d = cpml_update(P.p.psi[1], psi_index_y(g, 1, qy, x, z), P.p.coef[1] + (g.gy + y) * 3, d);
        div_y += d;
#pragma unroll
        for (int j = 0; j < 8; ++j) w[j] = q_yz[j];
        d = stencil8(w, invh);
        if (qz >= 0) d = cpml_update(P.p.psi[2], psi_index_z(g, 1, qz, x, y), P.p.coef[2] + z * 3, d);
        div_y += d;
        // vz: D-x sxz, D-y syz, D+z szz (half).
        d = dglob<false>(sxz, i, 1, invh);
        if (qx >= 0) d = cpml_update(P.p.psi[0], psi_index_x(g, 2, qx, y, z), P.p.coef[0] + x * 3, d);
        float div_z = 0.0f;
        div_z += d;
        d = dglob<false>(syz, i, g.px, invh);
        if (qy >= 0) d = cpml_update(P.p.psi[1], psi_index_y(g, 2, qy, x, z), P.p.coef[1] + y * 3, d);
        div_z += d;
#pragma unroll
        for (int j = 0; j < 8; ++j) w[j] = q_zz[j + 1];
        d = stencil8(w, invh);
        if (qz >= 0) d = cpml_update(P.p.psi[2], psi_index_z(g, 2, qz, x, y), P.p.coef[2] + (g.gz + z) * 3, d);
        div_z += d;

        const float* ib = P.p.inv_rho;
        const float a = ib[i];
        const float hx = face_buoyancy(a, ib[i + 1]);
        const float hy = face_buoyancy(a, ib[i + g.px]);
        const float hz = face_buoyancy(a, ib[i + g.plane]);
        float vx = P.p.f[0][i] + P.dt * hx * div_x;
        float vy = P.p.f[1][i] + P.dt * hy * div_y;
        float vz = P.p.f[2][i] + P.dt * hz * div_z;
        if (P.src.active) {
            const int r2 = source_r2c(P.src, x, y, z);
            if (r2 >= 0) {
                const float value = P.src.scale * P.src.w[r2] / P.src.norm;
                vx += value * P.src.dir[0] * hx;
                vy += value * P.src.dir[1] * hy;
                vz += value * P.src.dir[2] * hz;
            }
        }
        P.p.f[0][i] = vx; P.p.f[1][i] = vy; P.p.f[2][i] = vz;
        bad |= !isfinite(vx) || !isfinite(vy) || !isfinite(vz);
    }
    if (bad) atomicOr(P.p.flags, 1u);
}

template <int BY>
__global__ __launch_bounds__(64 * BY) void stress_column(DevStep P, TileMap M) {
    int bx, by, bz;
    if (!tile_of(M, bx, by, bz)) return;
    const DevGrid& g = P.g;
    const int x = 4 + bx * 64 + int(threadIdx.x), y = 4 + by * BY + int(threadIdx.y);
    if (x >= g.gx - 4 || y >= g.gy - 4) return;
    int z0, z1;
    chunk_of(M, g.gz, bz, z0, z1);
    const int qx = slab_q(x, g.gx, g.pml, false);
    const int qy = slab_q(y, g.gy, g.pml, false);
    const float invh = P.invh;
    const float* fvx = P.p.f[0];
    const float* fvy = P.p.f[1];
    const float* fvz = P.p.f[2];
    const long long col = ccell(g, x, y, 0);

    float q_vx[9], q_vy[9], q_vz[9];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const long long k = col + static_cast<long long>(z0 - 4 + j) * g.plane;
        q_vx[j + 1] = fvx[k];
        q_vy[j + 1] = fvy[k];
        q_vz[j + 1] = fvz[k];
    }
    bool bad = false;
    for (int z = z0; z < z1; ++z) {
        const long long i = col + static_cast<long long>(z) * g.plane;
        const long long lead = i + 4 * g.plane;
        shift(q_vx, fvx[lead]);
        shift(q_vy, fvy[lead]);
        shift(q_vz, fvz[lead]);
        const int qz = slab_q(z, g.gz, g.pml, g.free_surface != 0);
        float w[8];
        // vx: D-x (slot 3), D+y (half), D+z (half).
        float d00 = dglob<false>(fvx, i, 1, invh);
        if (qx >= 0) d00 = cpml_update(P.p.psi[0], psi_index_x(g, 3, qx, y, z), P.p.coef[0] + x * 3, d00);
        float d01 = dglob<true>(fvx, i, g.px, invh);
        if (qy >= 0) d01 = cpml_update(P.p.psi[1], psi_index_y(g, 3, qy, x, z), P.p.coef[1] + (g.gy + y) * 3, d01);
#pragma unroll
        for (int j = 0; j < 8; ++j) w[j] = q_vx[j + 1];
        float d02 = stencil8(w, invh);
        if (qz >= 0) d02 = cpml_update(P.p.psi[2], psi_index_z(g, 3, qz, x, y), P.p.coef[2] + (g.gz + z) * 3, d02);
        // vy: D+x (half), D-y, D+z (half).
        float d10 = dglob<true>(fvy, i, 1, invh);
        if (qx >= 0) d10 = cpml_update(P.p.psi[0], psi_index_x(g, 4, qx, y, z), P.p.coef[0] + (g.gx + x) * 3, d10);
        float d11 = dglob<false>(fvy, i, g.px, invh);
        if (qy >= 0) d11 = cpml_update(P.p.psi[1], psi_index_y(g, 4, qy, x, z), P.p.coef[1] + y * 3, d11);
#pragma unroll
        for (int j = 0; j < 8; ++j) w[j] = q_vy[j + 1];
        float d12 = stencil8(w, invh);
        if (qz >= 0) d12 = cpml_update(P.p.psi[2], psi_index_z(g, 4, qz, x, y), P.p.coef[2] + (g.gz + z) * 3, d12);
        // vz: D+x (half), D+y (half), D-z.
        float d20 = dglob<true>(fvz, i, 1, invh);
        if (qx >= 0) d20 = cpml_update(P.p.psi[0], psi_index_x(g, 5, qx, y, z), P.p.coef[0] + (g.gx + x) * 3, d20);
        float d21 = dglob<true>(fvz, i, g.px, invh);
        if (qy >= 0) d21 = cpml_update(P.p.psi[1], psi_index_y(g, 5, qy, x, z), P.p.coef[1] + (g.gy + y) * 3, d21);
#pragma unroll
        for (int j = 0; j < 8; ++j) w[j] = q_vz[j];
        float d22 = stencil8(w, invh);
        if (qz >= 0) d22 = cpml_update(P.p.psi[2], psi_index_z(g, 5, qz, x, y), P.p.coef[2] + z * 3, d22);

        const float* m = P.p.mu;
        const float lam = P.p.lambda[i], mu = m[i];
        const bool surface = g.free_surface && z == 4;
        if (surface) d22 = -lam / (lam + 2 * mu) * (d00 + d11);
        const float trace = d00 + d11 + d22;
        const long long sy = g.px, sz = g.plane;
        float sxx = P.p.f[3][i] + P.dt * (lam * trace + 2 * mu * d00);
        float syy = P.p.f[4][i] + P.dt * (lam * trace + 2 * mu * d11);
        float szz = surface ? 0.0f : P.p.f[5][i] + P.dt * (lam * trace + 2 * mu * d22);
        const float* r = P.p.inv_mu;
        const float r0 = r[i], rx = r[i + 1], ry = r[i + sy], rz = r[i + sz];
        const float sxy = P.p.f[6][i] + P.dt * shear_from_inverse(r0, rx, ry, r[i + 1 + sy]) * (d01 + d10);
        const float syz = P.p.f[7][i] + P.dt * shear_from_inverse(r0, ry, rz, r[i + sy + sz]) * (d12 + d21);
        const float sxz = P.p.f[8][i] + P.dt * shear_from_inverse(r0, rx, rz, r[i + 1 + sz]) * (d02 + d20);
        if (P.src.active) {
            const int r2 = source_r2c(P.src, x, y, z);
            if (r2 >= 0) {
                const float value = P.src.scale * P.src.w[r2] / P.src.norm;
                sxx += value; syy += value; szz += value;
            }
        }
        P.p.f[3][i] = sxx; P.p.f[4][i] = syy; P.p.f[5][i] = szz;
        P.p.f[6][i] = sxy; P.p.f[7][i] = syz; P.p.f[8][i] = sxz;
        bad |= !isfinite(sxx) || !isfinite(syy) || !isfinite(szz) || !isfinite(sxy) || !isfinite(syz) || !isfinite(sxz);
    }
    if (bad) atomicOr(P.p.flags, 1u);
}

hipError_t launch_column(bool velocity, const DevStep& P, int tile_y, int zchunk, hipStream_t s) {
    TileMap m;
    m.nbx = (P.g.gx - 8 + 63) / 64;
    m.nby = (P.g.gy - 8 + tile_y - 1) / tile_y;
    m.zchunk = zchunk;
    m.nbz = (P.g.gz - 8 + zchunk - 1) / zchunk;
    m.nblocks = m.nbx * m.nby * m.nbz;
    const int launched = (m.nblocks + 7) / 8 * 8;
    m.per_xcd = launched / 8;
    const dim3 grid(launched);
    switch (tile_y) {
    case 1:
        if (velocity) hipLaunchKernelGGL(velocity_column<1>, grid, dim3(64, 1), 0, s, P, m);
        else hipLaunchKernelGGL(stress_column<1>, grid, dim3(64, 1), 0, s, P, m);
        break;
    case 2:
        if (velocity) hipLaunchKernelGGL(velocity_column<2>, grid, dim3(64, 2), 0, s, P, m);
        else hipLaunchKernelGGL(stress_column<2>, grid, dim3(64, 2), 0, s, P, m);
        break;
    case 4:
        if (velocity) hipLaunchKernelGGL(velocity_column<4>, grid, dim3(64, 4), 0, s, P, m);
        else hipLaunchKernelGGL(stress_column<4>, grid, dim3(64, 4), 0, s, P, m);
        break;
    case 8:
        if (velocity) hipLaunchKernelGGL(velocity_column<8>, grid, dim3(64, 8), 0, s, P, m);
        else hipLaunchKernelGGL(stress_column<8>, grid, dim3(64, 8), 0, s, P, m);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace seis3d
