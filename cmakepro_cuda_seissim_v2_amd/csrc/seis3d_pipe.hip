// seis3d_pipe.hip - software-pipelined 2.5D velocity and stress half-steps.
//
// Same arithmetic, operation for operation, as the direct and stream kernels
// (velocityKernel3D / stressKernel3D, cuda/Cuda_Seis3D.cu:222-310), organised
// for latency hiding on CDNA4:
//  * a 64 x BY workgroup owns an x-y tile and marches along z through a chunk
//    of planes; wave w stages tile rows w, w+BY, ... with one coalesced
//    64-lane load per row plus an 8-lane tail for the x halo;
//  * the rows of lead plane L+1 are loaded into registers while plane L is
//    being computed, then written to the other half of a double-buffered LDS
//    tile, so each plane costs one workgroup barrier;
//  * x/y derivatives are CPML-filtered at the lead plane and queued (5 deep),
//    z derivatives come from 9-deep register queues of the tile's columns;
//  * the loads of the cell update (velocities or stresses, materials) are
//    issued before the prefetch so that waiting for them never waits on it.
#include "seis3d_device.hpp"
#include "seis3d_kernels.hpp"

namespace seis3d {

namespace {

__device__ __forceinline__ long long pcell(const DevGrid& g, int x, int y, int z) {
    return (static_cast<long long>(z) * g.gy + y) * g.px + x;
}

__device__ __forceinline__ int source_r2p(const DevSource& s, int x, int y, int z) {
    const int dx = x - s.x, dy = y - s.y, dz = z - s.z;
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return -1;
    return dx * dx + dy * dy + dz * dz;
}

// Rows of one staged tile held in registers between the load and the LDS
// write.  ROWS tile rows, W = 72 (x halo) or 64 columns; wave ty owns rows
// ty, ty+BY, ...
template <int BY, int ROWS, int W>
struct Rows {
    static constexpr int PER = (ROWS + BY - 1) / BY;
    float main[PER];
    float tail[PER];

    // Columns start at xs (x0 - 4 for W = 72, x0 for W = 64), rows at ys.
    __device__ __forceinline__ void load(const float* __restrict__ f, const DevGrid& g, int xs, int ys, int L,
                                         int tx, int ty) {
        const long long base = static_cast<long long>(L) * g.plane;
        const int xm = min(xs + tx, g.gx - 1);
        const int xt = min(xs + 64 + (tx & 7), g.gx - 1);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int r = ty + k * BY;
            if (ROWS % BY == 0 || r < ROWS) {
                const long long row = base + static_cast<long long>(min(ys + r, g.gy - 1)) * g.px;
                main[k] = f[row + xm];
                if (W == 72) tail[k] = f[row + xt];
            }
        }
    }

    __device__ __forceinline__ void store(float* __restrict__ lds, int tx, int ty) const {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int r = ty + k * BY;
            if (ROWS % BY == 0 || r < ROWS) {
                lds[r * W + tx] = main[k];
                if (W == 72 && tx < 8) lds[r * W + 64 + tx] = tail[k];
            }
        }
    }
};

template <int BY>
struct VelocityStage {
    Rows<BY, BY, 72> xx;
    Rows<BY, BY + 8, 72> xy;
    Rows<BY, BY + 8, 64> yy;
    Rows<BY, BY, 72> xz;
    Rows<BY, BY + 8, 64> yz;
    float zz;
};

template <int BY>
struct VelocityLds {
    float xx[BY * 72];
    float xy[(BY + 8) * 72];
    float yy[(BY + 8) * 64];
    float xz[BY * 72];
    float yz[(BY + 8) * 64];
};

}  // namespace

// Fused velocity half-step (velocityKernel3D + force source epilogue).
template <int BY>
__global__ __launch_bounds__(64 * BY) void velocity_pipe(DevStep P, TileMap M) {
    __shared__ VelocityLds<BY> lds[2];
    int bx, by, bz;
    if (!tile_of(M, bx, by, bz)) return;
    const DevGrid& g = P.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x0 = 4 + bx * 64, y0 = 4 + by * BY;
    const int x = x0 + tx, y = y0 + ty;
    int z0, z1;
    chunk_of(M, g.gz, bz, z0, z1);
    const bool active = x < g.gx - 4 && y < g.gy - 4;
    const int xc = min(x, g.gx - 1), yc = min(y, g.gy - 1);
    const int qx = slab_q(x, g.gx, g.pml, false);
    const int qy = slab_q(y, g.gy, g.pml, false);
    const float invh = P.invh;
    const float* sxx = P.p.f[3];
    const float* syy = P.p.f[4];
    const float* szz = P.p.f[5];
    const float* sxy = P.p.f[6];
    const float* syz = P.p.f[7];
    const float* sxz = P.p.f[8];

    float q_xz[9], q_yz[9], q_zz[9];
    float p_vx[5], p_vy[5], p_vz[5];
#pragma unroll
    for (int j = 0; j < 9; ++j) { q_xz[j] = 0.0f; q_yz[j] = 0.0f; q_zz[j] = 0.0f; }
#pragma unroll
    for (int j = 0; j < 5; ++j) { p_vx[j] = 0.0f; p_vy[j] = 0.0f; p_vz[j] = 0.0f; }

    VelocityStage<BY> R;
    auto prefetch = [&](int L) {
        R.xx.load(sxx, g, x0 - 4, y0, L, tx, ty);
        R.xy.load(sxy, g, x0 - 4, y0 - 4, L, tx, ty);
        R.yy.load(syy, g, x0, y0 - 4, L, tx, ty);
        R.xz.load(sxz, g, x0 - 4, y0, L, tx, ty);
        R.yz.load(syz, g, x0, y0 - 4, L, tx, ty);
        R.zz = szz[pcell(g, xc, yc, L)];
    };
    prefetch(z0 - 4);
    bool bad = false;
    int buf = 0;
    for (int L = z0 - 4; L < z1 + 4; ++L, buf ^= 1) {
        VelocityLds<BY>& s = lds[buf];
        R.xx.store(s.xx, tx, ty);
        R.xy.store(s.xy, tx, ty);
        R.yy.store(s.yy, tx, ty);
        R.xz.store(s.xz, tx, ty);
        R.yz.store(s.yz, tx, ty);
        const float lead_zz = R.zz;
        // Cell-update loads for z = L - 4 first, then the next plane's rows.
        const int z = L - 4;
        const bool update = active && z >= z0;
        const long long i = pcell(g, xc, yc, max(z, 0));
        float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f, b0 = 0.0f, bxn = 0.0f, byn = 0.0f, bzn = 0.0f;
        if (update) {
            v0 = P.p.f[0][i]; v1 = P.p.f[1][i]; v2 = P.p.f[2][i];
            const float* ib = P.p.inv_rho;
            b0 = ib[i]; bxn = ib[i + 1]; byn = ib[i + g.px]; bzn = ib[i + g.plane];
        }
        if (L + 1 < z1 + 4) prefetch(L + 1);
        __syncthreads();

        shift(q_xz, s.xz[ty * 72 + tx + 4]);
        shift(q_yz, s.yz[(ty + 4) * 64 + tx]);
        shift(q_zz, lead_zz);
        float pvx = 0.0f, pvy = 0.0f, pvz = 0.0f;
        if (active && L >= z0 && L < z1) {
            float w[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.xx[ty * 72 + tx + 1 + j];
            float d = stencil8(w, invh);
            if (qx >= 0) d = cpml_update(P.p.psi[0], psi_index_x(g, 0, qx, y, L), P.p.coef[0] + (g.gx + x) * 3, d);
            pvx += d;
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.xy[(ty + j) * 72 + tx + 4];
            d = stencil8(w, invh);
            if (qy >= 0) d = cpml_update(P.p.psi[1], psi_index_y(g, 0, qy, x, L), P.p.coef[1] + y * 3, d);
            pvx += d;
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.xy[(ty + 4) * 72 + tx + j];
            d = stencil8(w, invh);
            if (qx >= 0) d = cpml_update(P.p.psi[0], psi_index_x(g, 1, qx, y, L), P.p.coef[0] + x * 3, d);
            pvy += d;
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.yy[(ty + 1 + j) * 64 + tx];
            d = stencil8(w, invh);
            if (qy >= 0) d = cpml_update(P.p.psi[1], psi_index_y(g, 1, qy, x, L), P.p.coef[1] + (g.gy + y) * 3, d);
            pvy += d;
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.xz[ty * 72 + tx + j];
            d = stencil8(w, invh);
            if (qx >= 0) d = cpml_update(P.p.psi[0], psi_index_x(g, 2, qx, y, L), P.p.coef[0] + x * 3, d);
            pvz += d;
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.yz[(ty + j) * 64 + tx];
            d = stencil8(w, invh);
            if (qy >= 0) d = cpml_update(P.p.psi[1], psi_index_y(g, 2, qy, x, L), P.p.coef[1] + y * 3, d);
            pvz += d;
        }
        shift(p_vx, pvx);
        shift(p_vy, pvy);
        shift(p_vz, pvz);

        if (update) {
            const int qz = slab_q(z, g.gz, g.pml, g.free_surface != 0);
            const float hx = face_buoyancy(b0, bxn);
            const float hy = face_buoyancy(b0, byn);
            const float hz = face_buoyancy(b0, bzn);
            float w[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = q_xz[j];
            float dz = stencil8(w, invh);
            if (qz >= 0) dz = cpml_update(P.p.psi[2], psi_index_z(g, 0, qz, x, y), P.p.coef[2] + z * 3, dz);
            const float div_x = p_vx[0] + dz;
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = q_yz[j];
            dz = stencil8(w, invh);
            if (qz >= 0) dz = cpml_update(P.p.psi[2], psi_index_z(g, 1, qz, x, y), P.p.coef[2] + z * 3, dz);
            const float div_y = p_vy[0] + dz;
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = q_zz[j + 1];
            dz = stencil8(w, invh);
            if (qz >= 0) dz = cpml_update(P.p.psi[2], psi_index_z(g, 2, qz, x, y), P.p.coef[2] + (g.gz + z) * 3, dz);
            const float div_z = p_vz[0] + dz;
            float vx = v0 + P.dt * hx * div_x;
            float vy = v1 + P.dt * hy * div_y;
            float vz = v2 + P.dt * hz * div_z;
            if (P.src.active) {
                const int r2 = source_r2p(P.src, x, y, z);
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
    }
    if (bad) atomicOr(P.p.flags, 1u);
}

namespace {
template <int BY>
struct StressStage {
    Rows<BY, BY + 8, 72> vx, vy, vz;
};
template <int BY>
struct StressLds {
    float vx[(BY + 8) * 72];
    float vy[(BY + 8) * 72];
    float vz[(BY + 8) * 72];
};
}  // namespace

// Fused stress half-step (stressKernel3D + explosion source epilogue).
template <int BY>
__global__ __launch_bounds__(64 * BY) void stress_pipe(DevStep P, TileMap M) {
    __shared__ StressLds<BY> lds[2];
    int bx, by, bz;
    if (!tile_of(M, bx, by, bz)) return;
    const DevGrid& g = P.g;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x0 = 4 + bx * 64, y0 = 4 + by * BY;
    const int x = x0 + tx, y = y0 + ty;
    int z0, z1;
    chunk_of(M, g.gz, bz, z0, z1);
    const bool active = x < g.gx - 4 && y < g.gy - 4;
    const int xc = min(x, g.gx - 1), yc = min(y, g.gy - 1);
    const int qx = slab_q(x, g.gx, g.pml, false);
    const int qy = slab_q(y, g.gy, g.pml, false);
    const float invh = P.invh;

    float q_vx[9], q_vy[9], q_vz[9];
    float p00[5], p11[5], p01[5], p20[5], p21[5];
#pragma unroll
    for (int j = 0; j < 9; ++j) { q_vx[j] = 0.0f; q_vy[j] = 0.0f; q_vz[j] = 0.0f; }
#pragma unroll
    for (int j = 0; j < 5; ++j) { p00[j] = 0.0f; p11[j] = 0.0f; p01[j] = 0.0f; p20[j] = 0.0f; p21[j] = 0.0f; }

    StressStage<BY> R;
    auto prefetch = [&](int L) {
        R.vx.load(P.p.f[0], g, x0 - 4, y0 - 4, L, tx, ty);
        R.vy.load(P.p.f[1], g, x0 - 4, y0 - 4, L, tx, ty);
        R.vz.load(P.p.f[2], g, x0 - 4, y0 - 4, L, tx, ty);
    };
    prefetch(z0 - 4);
    bool bad = false;
    int buf = 0;
    for (int L = z0 - 4; L < z1 + 4; ++L, buf ^= 1) {
        StressLds<BY>& s = lds[buf];
        R.vx.store(s.vx, tx, ty);
        R.vy.store(s.vy, tx, ty);
        R.vz.store(s.vz, tx, ty);
        const int z = L - 4;
        const bool update = active && z >= z0;
        const long long i = pcell(g, xc, yc, max(z, 0));
        float o[6] = {0, 0, 0, 0, 0, 0};
        float lam = 0.0f, m0 = 0.0f, r0 = 0.0f, mx = 0.0f, my = 0.0f, mz = 0.0f, mxy = 0.0f, myz = 0.0f, mxz = 0.0f;
        if (update) {
#pragma unroll
            for (int k = 0; k < 6; ++k) o[k] = P.p.f[3 + k][i];
            const float* m = P.p.mu;
            lam = P.p.lambda[i];
            const float* r = P.p.inv_mu;
            m0 = m[i];
            r0 = r[i]; mx = r[i + 1]; my = r[i + g.px]; mz = r[i + g.plane];
            mxy = r[i + 1 + g.px]; myz = r[i + g.px + g.plane]; mxz = r[i + 1 + g.plane];
        }
        if (L + 1 < z1 + 4) prefetch(L + 1);
        __syncthreads();

        const int c = (ty + 4) * 72 + tx + 4;
        shift(q_vx, s.vx[c]);
        shift(q_vy, s.vy[c]);
        shift(q_vz, s.vz[c]);
        float d00 = 0.0f, d11 = 0.0f, s01 = 0.0f, d20 = 0.0f, d21 = 0.0f;
        if (active && L >= z0 && L < z1) {
            float w[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.vx[(ty + 4) * 72 + tx + j];
            d00 = stencil8(w, invh);
            if (qx >= 0) d00 = cpml_update(P.p.psi[0], psi_index_x(g, 3, qx, y, L), P.p.coef[0] + x * 3, d00);
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.vx[(ty + 1 + j) * 72 + tx + 4];
            float d01 = stencil8(w, invh);
            if (qy >= 0) d01 = cpml_update(P.p.psi[1], psi_index_y(g, 3, qy, x, L), P.p.coef[1] + (g.gy + y) * 3, d01);
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.vy[(ty + 4) * 72 + tx + 1 + j];
            float d10 = stencil8(w, invh);
            if (qx >= 0) d10 = cpml_update(P.p.psi[0], psi_index_x(g, 4, qx, y, L), P.p.coef[0] + (g.gx + x) * 3, d10);
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.vy[(ty + j) * 72 + tx + 4];
            d11 = stencil8(w, invh);
            if (qy >= 0) d11 = cpml_update(P.p.psi[1], psi_index_y(g, 4, qy, x, L), P.p.coef[1] + y * 3, d11);
            s01 = d01 + d10;
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.vz[(ty + 4) * 72 + tx + 1 + j];
            d20 = stencil8(w, invh);
            if (qx >= 0) d20 = cpml_update(P.p.psi[0], psi_index_x(g, 5, qx, y, L), P.p.coef[0] + (g.gx + x) * 3, d20);
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.vz[(ty + 1 + j) * 72 + tx + 4];
            d21 = stencil8(w, invh);
            if (qy >= 0) d21 = cpml_update(P.p.psi[1], psi_index_y(g, 5, qy, x, L), P.p.coef[1] + (g.gy + y) * 3, d21);
        }
        shift(p00, d00); shift(p11, d11); shift(p01, s01); shift(p20, d20); shift(p21, d21);

        if (update) {
            const int qz = slab_q(z, g.gz, g.pml, g.free_surface != 0);
            float w[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = q_vx[j + 1];
            float d02 = stencil8(w, invh);
            if (qz >= 0) d02 = cpml_update(P.p.psi[2], psi_index_z(g, 3, qz, x, y), P.p.coef[2] + (g.gz + z) * 3, d02);
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = q_vy[j + 1];
            float d12 = stencil8(w, invh);
            if (qz >= 0) d12 = cpml_update(P.p.psi[2], psi_index_z(g, 4, qz, x, y), P.p.coef[2] + (g.gz + z) * 3, d12);
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = q_vz[j];
            float d22 = stencil8(w, invh);
            if (qz >= 0) d22 = cpml_update(P.p.psi[2], psi_index_z(g, 5, qz, x, y), P.p.coef[2] + z * 3, d22);

            const bool surface = g.free_surface && z == 4;
            const float e00 = p00[0], e11 = p11[0];
            if (surface) d22 = -lam / (lam + 2 * m0) * (e00 + e11);
            const float trace = e00 + e11 + d22;
            float sxx = o[0] + P.dt * (lam * trace + 2 * m0 * e00);
            float syy = o[1] + P.dt * (lam * trace + 2 * m0 * e11);
            float szz = surface ? 0.0f : o[2] + P.dt * (lam * trace + 2 * m0 * d22);
            const float sxy = o[3] + P.dt * shear_from_inverse(r0, mx, my, mxy) * p01[0];
            const float syz = o[4] + P.dt * shear_from_inverse(r0, my, mz, myz) * (d12 + p21[0]);
            const float sxz = o[5] + P.dt * shear_from_inverse(r0, mx, mz, mxz) * (d02 + p20[0]);
            if (P.src.active) {
                const int r2 = source_r2p(P.src, x, y, z);
                if (r2 >= 0) {
                    const float value = P.src.scale * P.src.w[r2] / P.src.norm;
                    sxx += value; syy += value; szz += value;
                }
            }
            P.p.f[3][i] = sxx; P.p.f[4][i] = syy; P.p.f[5][i] = szz;
            P.p.f[6][i] = sxy; P.p.f[7][i] = syz; P.p.f[8][i] = sxz;
            bad |= !isfinite(sxx) || !isfinite(syy) || !isfinite(szz) || !isfinite(sxy) || !isfinite(syz) ||
                   !isfinite(sxz);
        }
    }
    if (bad) atomicOr(P.p.flags, 1u);
}

hipError_t launch_pipe(bool velocity, const DevStep& P, int tile_y, int zchunk, hipStream_t s) {
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
    case 4:
        if (velocity) hipLaunchKernelGGL(velocity_pipe<4>, grid, dim3(64, 4), 0, s, P, m);
        else hipLaunchKernelGGL(stress_pipe<4>, grid, dim3(64, 4), 0, s, P, m);
        break;
    case 8:
        if (velocity) hipLaunchKernelGGL(velocity_pipe<8>, grid, dim3(64, 8), 0, s, P, m);
        else hipLaunchKernelGGL(stress_pipe<8>, grid, dim3(64, 8), 0, s, P, m);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace seis3d
