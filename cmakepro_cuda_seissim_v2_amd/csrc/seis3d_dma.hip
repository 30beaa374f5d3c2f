// seis3d_dma.hip - 2.5D velocity and stress half-steps staged by direct
// global-to-LDS loads (gfx950 global_load_lds_dword).
//
// Same arithmetic, operation for operation, as velocityKernel3D /
// stressKernel3D (cuda/Cuda_Seis3D.cu:222-310).  A 64 x BY workgroup owns an
// x-y tile and marches along z through a chunk of planes:
//  * every wavefield plane that needs an x or y derivative is copied from HBM
//    straight into LDS (one 64-lane row per instruction plus an 8-lane tail
//    for the x halo), with no VGPRs holding the data in flight;
//  * the LDS ring has three plane buffers; the copy of plane k+2 is the last
//    vector-memory instruction of iteration k, so at the top of iteration k+1
//    "at most N_DMA outstanding" (N_DMA = copies per plane per wave) proves
//    that plane k+1 has landed for this wave, and the workgroup barrier that
//    follows proves it for every wave and frees the oldest buffer;
//  * the per-cell operands (the fields being updated, materials, CPML
//    memories) are loaded into registers one plane ahead;
//  * x/y derivatives are CPML-filtered at the lead plane and queued (5 deep),
//    z derivatives come from 9-deep register queues of the tile's columns.
// The loads are explicit waits rather than compiler-tracked, so every wait
// is conservative: a wave may wait for more than it needs, never less.
#include <algorithm>
#include <cmath>
#include <vector>

#include "seis3d_device.hpp"
#include "seis3d_kernels.hpp"

namespace seis3d {

namespace {

template <class G>
__device__ __forceinline__ long long dcell(const G& g, int x, int y, int z) {
    return (static_cast<long long>(z) * g.gy + y) * g.px + x;
}

template <class S>
__device__ __forceinline__ float source_weight(const S& s, int r2) {
    return r2 == 0 ? s.w[0] : r2 == 1 ? s.w[1] : r2 == 2 ? s.w[2] : s.w[3];
}

template <class S>
__device__ __forceinline__ int source_r2d(const S& s, int x, int y, int z) {
    const int dx = x - s.x, dy = y - s.y, dz = z - s.z;
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return -1;
    return dx * dx + dy * dy + dz * dz;
}

// Wait until at most N vector-memory operations of this wave are outstanding
// (gfx9 s_waitcnt: vmcnt in bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
template <int N>
__device__ __forceinline__ void wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

// Every array the kernels address lives in global memory.  Parameters passed as a kernel argument
// carry that knowledge already; pointers read from a device table (the work-queue kernels, TABLE)
// are retyped as global here, so the tile bodies still use global (not flat) instructions.
template <typename T>
using global_ptr = __attribute__((address_space(1))) T*;
template <bool TABLE, typename T>
__device__ __forceinline__ auto gptr(T* p) {
    if constexpr (TABLE) return (global_ptr<T>)p;
    else return p;
}

// Lane l of the wave copies src (its own address) to lds_row[l].
template <class F>
__device__ __forceinline__ void copy_row(F src, float* lds_row) {
    __builtin_amdgcn_global_load_lds(src, lds_row, 4, 0, 0);
}

// Copy ROWS rows of one plane into an LDS tile of width W (72: x halo,
// starting at column x0 - 4; 64: starting at the given column) whose rows are
// PITCH floats apart.  Wave ty copies rows ty, ty + BY, ...  A 72-wide row
// takes two full-wave copies, columns 0..63 and 8..71: the overlap is written
// twice with identical values, and no copy depends on the exec mask, so every
// wave issues exactly COPIES instructions.
template <int BY, int ROWS, int W, int PITCH = W>
struct TileCopy {
    static constexpr int PER = (ROWS + BY - 1) / BY;
    static constexpr int COPIES = PER * (W == 72 ? 2 : 1);
    static_assert(ROWS % BY == 0, "rows per wave must be exact");

    template <class F, class G>   // const float*, or the same pointer typed as global memory
    __device__ __forceinline__ static void issue(float* lds, F f, const G& g, int xs, int ys, int L, int tx,
                                                 int ty) {
        const long long base = static_cast<long long>(L) * g.plane;
        const int xm = min(xs + tx, g.gx - 1);
        const int xt = min(xs + 8 + tx, g.gx - 1);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int r = ty + k * BY;
            const long long row = base + static_cast<long long>(min(ys + r, g.gy - 1)) * g.px;
            copy_row(f + row + xm, lds + r * PITCH);
            if (W == 72) copy_row(f + row + xt, lds + r * PITCH + 8);
        }
    }
};

// Wait until the copies of plane `next` (the newest ones in flight) are the
// only vector-memory operations this wave may still have outstanding.  Planes
// inside the chunk carry FULL copies per wave, halo planes LITE ones.
template <int FULL, int LITE>
__device__ __forceinline__ void wait_for_ring(bool next_full) {
    if (next_full) wait_vm<FULL>();
    else wait_vm<LITE>();
}

// Arrays a half-step writes: its inputs for the in-place step, the output arrays for
// an out-of-place step.  The choice is a template parameter, so the in-place kernels
// carry no second set of pointers (and no extra registers for them).
template <bool OOP, bool TABLE, class Step>
__device__ __forceinline__ auto out_field(const Step& P, int c) {
    return gptr<TABLE>(OOP ? P.p.out[c] : P.p.f[c]);
}
template <bool OOP, bool TABLE, class Step>
__device__ __forceinline__ auto out_psi(const Step& P, int a) {
    return gptr<TABLE>(OOP ? P.p.psi_out[a] : P.p.psi[a]);
}

template <int BY>
struct VelTiles {
    float xx[BY * 72];
    float xy[(BY + 8) * 72];
    float yy[(BY + 8) * 64];
    float xz[BY * 72];
    float yz[(BY + 8) * 64];
};

template <int BY>
constexpr int vel_copies() {
    return TileCopy<BY, BY, 72>::COPIES + TileCopy<BY, BY + 8, 72>::COPIES + TileCopy<BY, BY + 8, 64>::COPIES +
           TileCopy<BY, BY, 72>::COPIES + TileCopy<BY, BY + 8, 64>::COPIES;
}

// Per-cell operands of one velocity iteration (lead plane k, update plane k-4).
struct VelOperands {
    float zz;          // szz at the lead plane, own column
    float v[3];        // velocities at the update plane
    float b[4];        // 1/rho at i, i+x, i+y, i+z
    float pz[3];       // z-slab CPML memories at the update plane
};

constexpr int kMaxChunk = 128;   // planes per tile (z table capacity)


// Copy the tile's slices of the three CPML tables ([half][n][a,b,1/kappa])
// into LDS as [half*3 + k][index]; entries outside a slab are never read.
template <bool TABLE, class Step>
__device__ __forceinline__ void fill_tables(const Step& P, int x0, int y0, int z0, int z1, int by_rows,
                                            float* cx, float* cyt, float* czt) {
    const auto& g = P.g;
    if (g.pml == 0) return;
    const int tid = threadIdx.y * 64 + threadIdx.x, nt = 64 * by_rows;
    for (int k = tid; k < 6 * 64; k += nt) {
        const int j = k / 64, c = min(x0 + k % 64, g.gx - 1);
        cx[k] = gptr<TABLE>(P.p.coef[0])[((j / 3) * g.gx + c) * 3 + j % 3];
    }
    for (int k = tid; k < 6 * by_rows; k += nt) {
        const int j = k / by_rows, c = min(y0 + k % by_rows, g.gy - 1);
        cyt[k] = gptr<TABLE>(P.p.coef[1])[((j / 3) * g.gy + c) * 3 + j % 3];
    }
    const int nz = z1 - z0;
    for (int k = tid; k < 6 * nz; k += nt) {
        const int j = k / nz, c = z0 + k % nz;
        czt[j * kMaxChunk + k % nz] = gptr<TABLE>(P.p.coef[2])[((j / 3) * g.gz + c) * 3 + j % 3];
    }
}

}  // namespace

// One workgroup's tile (bx, by) and z-chunk bz of the velocity half-step.
// LDS of the velocity half-step: the plane ring and the tile's CPML tables.
template <int BY>
struct VelShared {
    VelTiles<BY> ring[3];
    float cx[6 * 64];          // x CPML table: [half*3 + k][lane]
    float cyt[6 * BY];         // y CPML table: [half*3 + k][row]
    float czt[6 * kMaxChunk];  // z CPML table: [half*3 + k][plane - z0]
};

template <int BY, bool OOP, bool TABLE = false, class Step = DevStep>
__device__ __forceinline__ void velocity_tile(VelShared<BY>& sh, const Step& P, const TileMap& M, int bx, int by,
                                              int bz) {
    VelTiles<BY>* ring = sh.ring;
    float* cx = sh.cx;
    float* cyt = sh.cyt;
    float* czt = sh.czt;
    const auto& g = P.g;
    const int tx = threadIdx.x;
    const int ty = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int x0 = 4 + bx * 64, y0 = 4 + by * BY;
    const int x = x0 + tx, y = y0 + ty;
    int z0, z1;
    chunk_of(M, g.gz, bz, z0, z1);
    const int k0 = z0 - 4, last = z1 + 3;
    const bool active = x < g.gx - 4 && y < g.gy - 4;
    const int xc = min(x, g.gx - 1), yc = min(y, g.gy - 1);
    const int qx = slab_q(x, g.gx, g.pml, false);
    const int qy = slab_q(y, g.gy, g.pml, false);
    const float invh = P.invh;
    const auto sxx = gptr<TABLE>(P.p.f[3]);
    const auto syy = gptr<TABLE>(P.p.f[4]);
    const auto szz = gptr<TABLE>(P.p.f[5]);
    const auto sxy = gptr<TABLE>(P.p.f[6]);
    const auto syz = gptr<TABLE>(P.p.f[7]);
    const auto sxz = gptr<TABLE>(P.p.f[8]);
    const auto ib = gptr<TABLE>(P.p.inv_rho);
    constexpr int NDMA = vel_copies<BY>();
    constexpr int NLITE = 2 * TileCopy<BY, BY, 64>::COPIES;

    // CPML coefficient tables of this tile (read from LDS: no vector-memory
    // waits inside the plane loop).
    fill_tables<TABLE>(P, x0, y0, z0, z1, BY, cx, cyt, czt);
    __syncthreads();

    // Planes z0..z1-1 need whole tiles (x/y derivatives); the halo planes
    // only feed the z queues, so they copy just the tile's own rows of sxz
    // and syz into the places the queues read.
    auto full_plane = [&](int slot) { return slot >= z0 && slot < z1; };
    auto copy_plane = [&](int slot, int L) {
        VelTiles<BY>& s = ring[(slot - k0) % 3];
        if (full_plane(slot)) {
            TileCopy<BY, BY, 72>::issue(s.xx, sxx, g, x0 - 4, y0, L, tx, ty);
            TileCopy<BY, BY + 8, 72>::issue(s.xy, sxy, g, x0 - 4, y0 - 4, L, tx, ty);
            TileCopy<BY, BY + 8, 64>::issue(s.yy, syy, g, x0, y0 - 4, L, tx, ty);
            TileCopy<BY, BY, 72>::issue(s.xz, sxz, g, x0 - 4, y0, L, tx, ty);
            TileCopy<BY, BY + 8, 64>::issue(s.yz, syz, g, x0, y0 - 4, L, tx, ty);
        } else {
            TileCopy<BY, BY, 64, 72>::issue(s.xz + 4, sxz, g, x0, y0, L, tx, ty);
            TileCopy<BY, BY, 64>::issue(s.yz + 4 * 64, syz, g, x0, y0, L, tx, ty);
        }
    };
    auto load_operands = [&](int k, VelOperands& u) {
        u.zz = szz[dcell(g, xc, yc, k)];
        const int z = k - 4;
        if (active && z >= z0) {
            const long long i = dcell(g, x, y, z);
#pragma unroll
            for (int c = 0; c < 3; ++c) u.v[c] = gptr<TABLE>(P.p.f[c])[i];
            u.b[0] = ib[i]; u.b[1] = ib[i + 1]; u.b[2] = ib[i + g.px]; u.b[3] = ib[i + g.plane];
            const int qz = slab_q(z, g.gz, g.pml, g.free_surface != 0);
            if (qz >= 0) {
#pragma unroll
                for (int s = 0; s < 3; ++s) u.pz[s] = gptr<TABLE>(P.p.psi[2])[psi_index_z(g, s, qz, x, y)];
            }
        }
    };

    float q_xz[9], q_yz[9], q_zz[9];     // columns k-8..k (update plane z = k-4 in the middle)
    float p_vx[4], p_vy[4], p_vz[4];     // CPML-filtered x+y sums of planes k-4..k-1
#pragma unroll
    for (int j = 0; j < 9; ++j) { q_xz[j] = 0.0f; q_yz[j] = 0.0f; q_zz[j] = 0.0f; }
#pragma unroll
    for (int j = 0; j < 4; ++j) { p_vx[j] = 0.0f; p_vy[j] = 0.0f; p_vz[j] = 0.0f; }

    // Two operand sets used alternately (the loop is unrolled by two), so the
    // prefetched registers are never copied while their loads are in flight.
    VelOperands opa{}, opb{};
    copy_plane(k0, k0);
    load_operands(k0, opa);
    copy_plane(k0 + 1, k0 + 1);
    bool bad = false;
    auto plane_step = [&](int k, const VelOperands& cur, VelOperands& nxt) {
        const VelTiles<BY>& s = ring[(k - k0) % 3];
        wait_for_ring<NDMA, NLITE>(full_plane(k + 1));
        __builtin_amdgcn_s_barrier();
        // The wave issues this iteration's loads at raised priority (s_setprio: SIMD issue
        // arbitration among resident waves; 2 here, 3 for the ring copies below), so that the
        // memory pipeline is fed before the other waves' arithmetic.
        __builtin_amdgcn_s_setprio(2);
        if (k + 1 <= last) load_operands(k + 1, nxt);
        // x/y CPML memories of plane k, used after the update below has run.
        float px[3], py[3];
        if (active && k >= z0 && k < z1) {
            if (qx >= 0) {
#pragma unroll
                for (int c = 0; c < 3; ++c) px[c] = gptr<TABLE>(P.p.psi[0])[psi_index_x(g, c, qx, y, k)];
            }
            if (qy >= 0) {
#pragma unroll
                for (int c = 0; c < 3; ++c) py[c] = gptr<TABLE>(P.p.psi[1])[psi_index_y(g, c, qy, x, k)];
            }
        }
        __builtin_amdgcn_s_setprio(0);

        shift(q_xz, s.xz[ty * 72 + tx + 4]);
        shift(q_yz, s.yz[(ty + 4) * 64 + tx]);
        shift(q_zz, cur.zz);
        // Update plane z = k - 4 first (its x/y sums are p_*[0]) so its stores
        // drain while plane k's x/y derivatives are computed.
        const int z = k - 4;
        if (active && z >= z0) {
            const int qz = slab_q(z, g.gz, g.pml, g.free_surface != 0);
            const int zt = z - z0;
            const float c0[3] = {czt[0 * kMaxChunk + zt], czt[1 * kMaxChunk + zt], czt[2 * kMaxChunk + zt]};
            const float c1[3] = {czt[3 * kMaxChunk + zt], czt[4 * kMaxChunk + zt], czt[5 * kMaxChunk + zt]};
            const float hx = face_buoyancy(cur.b[0], cur.b[1]);
            const float hy = face_buoyancy(cur.b[0], cur.b[2]);
            const float hz = face_buoyancy(cur.b[0], cur.b[3]);
            float w[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = q_xz[j];
            float dz = stencil8(w, invh);
            if (qz >= 0) {
                const float m = c0[1] * cur.pz[0] + c0[0] * dz;
                out_psi<OOP, TABLE>(P, 2)[psi_index_z(g, 0, qz, x, y)] = m;
                dz = c0[2] * dz + m;
            }
            const float div_x = p_vx[0] + dz;
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = q_yz[j];
            dz = stencil8(w, invh);
            if (qz >= 0) {
                const float m = c0[1] * cur.pz[1] + c0[0] * dz;
                out_psi<OOP, TABLE>(P, 2)[psi_index_z(g, 1, qz, x, y)] = m;
                dz = c0[2] * dz + m;
            }
            const float div_y = p_vy[0] + dz;
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = q_zz[j + 1];
            dz = stencil8(w, invh);
            if (qz >= 0) {
                const float m = c1[1] * cur.pz[2] + c1[0] * dz;
                out_psi<OOP, TABLE>(P, 2)[psi_index_z(g, 2, qz, x, y)] = m;
                dz = c1[2] * dz + m;
            }
            const float div_z = p_vz[0] + dz;
            float vx = cur.v[0] + P.dt * hx * div_x;
            float vy = cur.v[1] + P.dt * hy * div_y;
            float vz = cur.v[2] + P.dt * hz * div_z;
            if (P.src.active) {
                const int r2 = source_r2d(P.src, x, y, z);
                if (r2 >= 0) {
                    const float value = P.src.scale * source_weight(P.src, r2) / P.src.norm;
                    vx += value * P.src.dir[0] * hx;
                    vy += value * P.src.dir[1] * hy;
                    vz += value * P.src.dir[2] * hz;
                }
            }
            const long long i = dcell(g, x, y, z);
            out_field<OOP, TABLE>(P, 0)[i] = vx; out_field<OOP, TABLE>(P, 1)[i] = vy; out_field<OOP, TABLE>(P, 2)[i] = vz;
            bad |= !isfinite(vx) || !isfinite(vy) || !isfinite(vz);
        }
        float pvx = 0.0f, pvy = 0.0f, pvz = 0.0f;
        if (active && k >= z0 && k < z1) {
            float w[8];
            // vx: D+x sxx (half), D-y sxy.
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.xx[ty * 72 + tx + 1 + j];
            float d = stencil8(w, invh);
            if (qx >= 0) {
                const float m = cx[4 * 64 + tx] * px[0] + cx[3 * 64 + tx] * d;
                out_psi<OOP, TABLE>(P, 0)[psi_index_x(g, 0, qx, y, k)] = m;
                d = cx[5 * 64 + tx] * d + m;
            }
            pvx += d;
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.xy[(ty + j) * 72 + tx + 4];
            d = stencil8(w, invh);
            if (qy >= 0) {
                const float m = cyt[1 * BY + ty] * py[0] + cyt[0 * BY + ty] * d;
                out_psi<OOP, TABLE>(P, 1)[psi_index_y(g, 0, qy, x, k)] = m;
                d = cyt[2 * BY + ty] * d + m;
            }
            pvx += d;
            // vy: D-x sxy, D+y syy (half).
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.xy[(ty + 4) * 72 + tx + j];
            d = stencil8(w, invh);
            if (qx >= 0) {
                const float m = cx[1 * 64 + tx] * px[1] + cx[0 * 64 + tx] * d;
                out_psi<OOP, TABLE>(P, 0)[psi_index_x(g, 1, qx, y, k)] = m;
                d = cx[2 * 64 + tx] * d + m;
            }
            pvy += d;
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.yy[(ty + 1 + j) * 64 + tx];
            d = stencil8(w, invh);
            if (qy >= 0) {
                const float m = cyt[4 * BY + ty] * py[1] + cyt[3 * BY + ty] * d;
                out_psi<OOP, TABLE>(P, 1)[psi_index_y(g, 1, qy, x, k)] = m;
                d = cyt[5 * BY + ty] * d + m;
            }
            pvy += d;
            // vz: D-x sxz, D-y syz.
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.xz[ty * 72 + tx + j];
            d = stencil8(w, invh);
            if (qx >= 0) {
                const float m = cx[1 * 64 + tx] * px[2] + cx[0 * 64 + tx] * d;
                out_psi<OOP, TABLE>(P, 0)[psi_index_x(g, 2, qx, y, k)] = m;
                d = cx[2 * 64 + tx] * d + m;
            }
            pvz += d;
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.yz[(ty + j) * 64 + tx];
            d = stencil8(w, invh);
            if (qy >= 0) {
                const float m = cyt[1 * BY + ty] * py[2] + cyt[0 * BY + ty] * d;
                out_psi<OOP, TABLE>(P, 1)[psi_index_y(g, 2, qy, x, k)] = m;
                d = cyt[2 * BY + ty] * d + m;
            }
            pvz += d;
        }
        shift(p_vx, pvx);
        shift(p_vy, pvy);
        shift(p_vz, pvz);
        // Last vector-memory instruction of the iteration: plane k+2 into the
        // buffer every wave finished reading before this iteration's barrier
        // (past the chunk end a clamped plane fills a buffer nobody reads).
        __builtin_amdgcn_s_setprio(3);
        copy_plane(k + 2, min(k + 2, last));
        __builtin_amdgcn_s_setprio(0);
    };
    int k = k0;
    for (; k + 1 <= last; k += 2) {
        plane_step(k, opa, opb);
        plane_step(k + 1, opb, opa);
    }
    if (k <= last) plane_step(k, opa, opb);
    wait_vm<0>();
    if (bad) atomicOr(P.p.flags, 1u);
}

namespace {

template <int BY>
struct StrTiles {
    float vx[(BY + 8) * 72];
    float vy[(BY + 8) * 72];
    float vz[(BY + 8) * 72];
};

template <int BY>
constexpr int str_copies() {
    return 3 * TileCopy<BY, BY + 8, 72>::COPIES;
}

struct StrOperands {
    float o[6];          // stresses at the update plane
    float lam, mu;
    float r[7];          // mu at i+x, i+y, i+z, i+x+y, i+y+z, i+x+z in r[1..6] (r[0] unused: mu)
    float pz[3];         // z-slab CPML memories at the update plane
};

}  // namespace

// One workgroup's tile (bx, by) and z-chunk bz of the stress half-step.
template <int BY>
struct StrShared {
    StrTiles<BY> ring[3];
    float cx[6 * 64];
    float cyt[6 * BY];
    float czt[6 * kMaxChunk];
};

template <int BY, bool OOP, bool TABLE = false, class Step = DevStep>
__device__ __forceinline__ void stress_tile(StrShared<BY>& sh, const Step& P, const TileMap& M, int bx, int by,
                                            int bz) {
    StrTiles<BY>* ring = sh.ring;
    float* cx = sh.cx;
    float* cyt = sh.cyt;
    float* czt = sh.czt;
    const auto& g = P.g;
    const int tx = threadIdx.x;
    const int ty = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int x0 = 4 + bx * 64, y0 = 4 + by * BY;
    const int x = x0 + tx, y = y0 + ty;
    int z0, z1;
    chunk_of(M, g.gz, bz, z0, z1);
    const int k0 = z0 - 4, last = z1 + 3;
    const bool active = x < g.gx - 4 && y < g.gy - 4;
    const int qx = slab_q(x, g.gx, g.pml, false);
    const int qy = slab_q(y, g.gy, g.pml, false);
    const float invh = P.invh;
    constexpr int NDMA = str_copies<BY>();
    constexpr int NLITE = 3 * TileCopy<BY, BY, 64>::COPIES;

    // CPML coefficient tables of this tile (read from LDS: no vector-memory
    // waits inside the plane loop).
    fill_tables<TABLE>(P, x0, y0, z0, z1, BY, cx, cyt, czt);
    __syncthreads();

    // Halo planes only feed the z queues: copy just the tile's own rows.
    auto full_plane = [&](int slot) { return slot >= z0 && slot < z1; };
    auto copy_plane = [&](int slot, int L) {
        StrTiles<BY>& s = ring[(slot - k0) % 3];
        if (full_plane(slot)) {
            TileCopy<BY, BY + 8, 72>::issue(s.vx, gptr<TABLE>(P.p.f[0]), g, x0 - 4, y0 - 4, L, tx, ty);
            TileCopy<BY, BY + 8, 72>::issue(s.vy, gptr<TABLE>(P.p.f[1]), g, x0 - 4, y0 - 4, L, tx, ty);
            TileCopy<BY, BY + 8, 72>::issue(s.vz, gptr<TABLE>(P.p.f[2]), g, x0 - 4, y0 - 4, L, tx, ty);
        } else {
            constexpr int own = 4 * 72 + 4;
            TileCopy<BY, BY, 64, 72>::issue(s.vx + own, gptr<TABLE>(P.p.f[0]), g, x0, y0, L, tx, ty);
            TileCopy<BY, BY, 64, 72>::issue(s.vy + own, gptr<TABLE>(P.p.f[1]), g, x0, y0, L, tx, ty);
            TileCopy<BY, BY, 64, 72>::issue(s.vz + own, gptr<TABLE>(P.p.f[2]), g, x0, y0, L, tx, ty);
        }
    };
    auto load_operands = [&](int k, StrOperands& u) {
        const int z = k - 4;
        if (active && z >= z0) {
            const long long i = dcell(g, x, y, z);
#pragma unroll
            for (int c = 0; c < 6; ++c) u.o[c] = gptr<TABLE>(P.p.f[3 + c])[i];
            u.lam = gptr<TABLE>(P.p.lambda)[i];
            u.mu = gptr<TABLE>(P.p.mu)[i];
            // Shear moduli around the cell: the harmonic means below divide by them
            // exactly as shearMu3D does (no 1/mu array to read).
            const auto r = gptr<TABLE>(P.p.mu);
            const long long sy = g.px, sz = g.plane;
            u.r[1] = r[i + 1]; u.r[2] = r[i + sy]; u.r[3] = r[i + sz];
            u.r[4] = r[i + 1 + sy]; u.r[5] = r[i + sy + sz]; u.r[6] = r[i + 1 + sz];
            const int qz = slab_q(z, g.gz, g.pml, g.free_surface != 0);
            if (qz >= 0) {
#pragma unroll
                for (int s = 0; s < 3; ++s) u.pz[s] = gptr<TABLE>(P.p.psi[2])[psi_index_z(g, 3 + s, qz, x, y)];
            }
        }
    };

    float q_vx[9], q_vy[9], q_vz[9];
    float p00[4], p11[4], p01[4], p20[4], p21[4];  // x/y terms of planes k-4..k-1
#pragma unroll
    for (int j = 0; j < 9; ++j) { q_vx[j] = 0.0f; q_vy[j] = 0.0f; q_vz[j] = 0.0f; }
#pragma unroll
    for (int j = 0; j < 4; ++j) { p00[j] = 0.0f; p11[j] = 0.0f; p01[j] = 0.0f; p20[j] = 0.0f; p21[j] = 0.0f; }

    // Two operand sets used alternately (the loop is unrolled by two), so the
    // prefetched registers are never copied while their loads are in flight.
    StrOperands opa{}, opb{};
    copy_plane(k0, k0);
    load_operands(k0, opa);
    copy_plane(k0 + 1, k0 + 1);
    bool bad = false;
    auto plane_step = [&](int k, const StrOperands& cur, StrOperands& nxt) {
        const StrTiles<BY>& s = ring[(k - k0) % 3];
        wait_for_ring<NDMA, NLITE>(full_plane(k + 1));
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_s_setprio(2);   // loads at raised issue priority, as in velocity_tile
        if (k + 1 <= last) load_operands(k + 1, nxt);
        // x/y CPML memories (slots 3-5) of plane k, used after the update below has run.
        float px[3], py[3];
        if (active && k >= z0 && k < z1) {
            if (qx >= 0) {
#pragma unroll
                for (int j = 0; j < 3; ++j) px[j] = gptr<TABLE>(P.p.psi[0])[psi_index_x(g, 3 + j, qx, y, k)];
            }
            if (qy >= 0) {
#pragma unroll
                for (int j = 0; j < 3; ++j) py[j] = gptr<TABLE>(P.p.psi[1])[psi_index_y(g, 3 + j, qy, x, k)];
            }
        }
        __builtin_amdgcn_s_setprio(0);
        const int c = (ty + 4) * 72 + tx + 4;
        shift(q_vx, s.vx[c]);
        shift(q_vy, s.vy[c]);
        shift(q_vz, s.vz[c]);
        // Update plane z = k - 4 first (its x/y terms are p*[0]) so its stores
        // drain while plane k's x/y derivatives are computed.
        const int z = k - 4;
        if (active && z >= z0) {
            const int qz = slab_q(z, g.gz, g.pml, g.free_surface != 0);
            const int zt = z - z0;
            const float c0[3] = {czt[0 * kMaxChunk + zt], czt[1 * kMaxChunk + zt], czt[2 * kMaxChunk + zt]};
            const float c1[3] = {czt[3 * kMaxChunk + zt], czt[4 * kMaxChunk + zt], czt[5 * kMaxChunk + zt]};
            float w[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = q_vx[j + 1];
            float d02 = stencil8(w, invh);
            if (qz >= 0) {
                const float m = c1[1] * cur.pz[0] + c1[0] * d02;
                out_psi<OOP, TABLE>(P, 2)[psi_index_z(g, 3, qz, x, y)] = m;
                d02 = c1[2] * d02 + m;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = q_vy[j + 1];
            float d12 = stencil8(w, invh);
            if (qz >= 0) {
                const float m = c1[1] * cur.pz[1] + c1[0] * d12;
                out_psi<OOP, TABLE>(P, 2)[psi_index_z(g, 4, qz, x, y)] = m;
                d12 = c1[2] * d12 + m;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = q_vz[j];
            float d22 = stencil8(w, invh);
            if (qz >= 0) {
                const float m = c0[1] * cur.pz[2] + c0[0] * d22;
                out_psi<OOP, TABLE>(P, 2)[psi_index_z(g, 5, qz, x, y)] = m;
                d22 = c0[2] * d22 + m;
            }
            const bool surface = g.free_surface && z == 4;
            const float lam = cur.lam, mu = cur.mu;
            const float e00 = p00[0], e11 = p11[0];
            if (surface) d22 = -lam / (lam + 2 * mu) * (e00 + e11);
            const float trace = e00 + e11 + d22;
            float sxx = cur.o[0] + P.dt * (lam * trace + 2 * mu * e00);
            float syy = cur.o[1] + P.dt * (lam * trace + 2 * mu * e11);
            float szz = surface ? 0.0f : cur.o[2] + P.dt * (lam * trace + 2 * mu * d22);
            // shearMu3D's reciprocals of the seven moduli around the cell, each divided once
            // (the three harmonic means share them; the same IEEE quotients, summed in the
            // same order).
            float inv[7];
            inv[0] = 1 / mu;
#pragma unroll
            for (int j = 1; j < 7; ++j) inv[j] = 1 / cur.r[j];
            auto shear = [&](int a, int b, int c) {
                return (mu > 0 && cur.r[a] > 0 && cur.r[b] > 0 && cur.r[c] > 0)
                           ? 4 / (inv[0] + inv[a] + inv[b] + inv[c]) : 0.0f;
            };
            const float sxy = cur.o[3] + P.dt * shear(1, 2, 4) * p01[0];
            const float syz = cur.o[4] + P.dt * shear(2, 3, 5) * (d12 + p21[0]);
            const float sxz = cur.o[5] + P.dt * shear(1, 3, 6) * (d02 + p20[0]);
            if (P.src.active) {
                const int r2 = source_r2d(P.src, x, y, z);
                if (r2 >= 0) {
                    const float value = P.src.scale * source_weight(P.src, r2) / P.src.norm;
                    sxx += value; syy += value; szz += value;
                }
            }
            const long long i = dcell(g, x, y, z);
            out_field<OOP, TABLE>(P, 3)[i] = sxx; out_field<OOP, TABLE>(P, 4)[i] = syy; out_field<OOP, TABLE>(P, 5)[i] = szz;
            out_field<OOP, TABLE>(P, 6)[i] = sxy; out_field<OOP, TABLE>(P, 7)[i] = syz; out_field<OOP, TABLE>(P, 8)[i] = sxz;
            bad |= !isfinite(sxx) || !isfinite(syy) || !isfinite(szz) || !isfinite(sxy) || !isfinite(syz) ||
                   !isfinite(sxz);
            // Free-surface ghost planes (stressGhostKernel3D, cuda/Cuda_Seis3D.cu:203-219) from
            // the planes this chunk just updated: normal stresses and sxy even about z = 4,
            // szz odd about z = 4, sxz and syz odd about z = 3.5.
            if (g.free_surface && z <= 8) {
                if (z >= 5) {
                    const long long gi = dcell(g, x, y, 8 - z);
                    out_field<OOP, TABLE>(P, 5)[gi] = -szz; out_field<OOP, TABLE>(P, 3)[gi] = sxx; out_field<OOP, TABLE>(P, 4)[gi] = syy; out_field<OOP, TABLE>(P, 6)[gi] = sxy;
                }
                if (z <= 7) {
                    const long long gi = dcell(g, x, y, 7 - z);
                    out_field<OOP, TABLE>(P, 8)[gi] = -sxz; out_field<OOP, TABLE>(P, 7)[gi] = -syz;
                }
            }
        }
        float d00 = 0.0f, d11 = 0.0f, s01 = 0.0f, d20 = 0.0f, d21 = 0.0f;
        if (active && k >= z0 && k < z1) {
            float w[8];
            // vx: D-x (slot 3, integer), D+y (slot 3, half).
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.vx[(ty + 4) * 72 + tx + j];
            d00 = stencil8(w, invh);
            if (qx >= 0) {
                const float m = cx[1 * 64 + tx] * px[0] + cx[0 * 64 + tx] * d00;
                out_psi<OOP, TABLE>(P, 0)[psi_index_x(g, 3, qx, y, k)] = m;
                d00 = cx[2 * 64 + tx] * d00 + m;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.vx[(ty + 1 + j) * 72 + tx + 4];
            float d01 = stencil8(w, invh);
            if (qy >= 0) {
                const float m = cyt[4 * BY + ty] * py[0] + cyt[3 * BY + ty] * d01;
                out_psi<OOP, TABLE>(P, 1)[psi_index_y(g, 3, qy, x, k)] = m;
                d01 = cyt[5 * BY + ty] * d01 + m;
            }
            // vy: D+x (slot 4, half), D-y (slot 4, integer).
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.vy[(ty + 4) * 72 + tx + 1 + j];
            float d10 = stencil8(w, invh);
            if (qx >= 0) {
                const float m = cx[4 * 64 + tx] * px[1] + cx[3 * 64 + tx] * d10;
                out_psi<OOP, TABLE>(P, 0)[psi_index_x(g, 4, qx, y, k)] = m;
                d10 = cx[5 * 64 + tx] * d10 + m;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.vy[(ty + j) * 72 + tx + 4];
            d11 = stencil8(w, invh);
            if (qy >= 0) {
                const float m = cyt[1 * BY + ty] * py[1] + cyt[0 * BY + ty] * d11;
                out_psi<OOP, TABLE>(P, 1)[psi_index_y(g, 4, qy, x, k)] = m;
                d11 = cyt[2 * BY + ty] * d11 + m;
            }
            s01 = d01 + d10;
            // vz: D+x (slot 5, half), D+y (slot 5, half).
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.vz[(ty + 4) * 72 + tx + 1 + j];
            d20 = stencil8(w, invh);
            if (qx >= 0) {
                const float m = cx[4 * 64 + tx] * px[2] + cx[3 * 64 + tx] * d20;
                out_psi<OOP, TABLE>(P, 0)[psi_index_x(g, 5, qx, y, k)] = m;
                d20 = cx[5 * 64 + tx] * d20 + m;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j) w[j] = s.vz[(ty + 1 + j) * 72 + tx + 4];
            d21 = stencil8(w, invh);
            if (qy >= 0) {
                const float m = cyt[4 * BY + ty] * py[2] + cyt[3 * BY + ty] * d21;
                out_psi<OOP, TABLE>(P, 1)[psi_index_y(g, 5, qy, x, k)] = m;
                d21 = cyt[5 * BY + ty] * d21 + m;
            }
        }
        shift(p00, d00); shift(p11, d11); shift(p01, s01); shift(p20, d20); shift(p21, d21);
        __builtin_amdgcn_s_setprio(3);
        copy_plane(k + 2, min(k + 2, last));
        __builtin_amdgcn_s_setprio(0);
    };
    int k = k0;
    for (; k + 1 <= last; k += 2) {
        plane_step(k, opa, opb);
        plane_step(k + 1, opb, opa);
    }
    if (k <= last) plane_step(k, opa, opb);
    wait_vm<0>();
    if (bad) atomicOr(P.p.flags, 1u);
}

// Fused velocity half-step (velocityKernel3D + force source epilogue).
template <int BY, bool OOP>
__global__ __launch_bounds__(64 * BY) __attribute__((amdgpu_waves_per_eu(4))) void velocity_dma(DevStep P, TileMap M) {
    __shared__ VelShared<BY> sh;
    int bx, by, bz;
    if (!tile_of(M, bx, by, bz)) return;
    velocity_tile<BY, OOP>(sh, P, M, bx, by, bz);
}

// Fused stress half-step (stressKernel3D + explosion source epilogue + the free surface's
// stress ghost planes).
template <int BY, bool OOP>
__global__ __launch_bounds__(64 * BY) void stress_dma(DevStep P, TileMap M) {
    __shared__ StrShared<BY> sh;
    int bx, by, bz;
    if (!tile_of(M, bx, by, bz)) return;
    stress_tile<BY, OOP>(sh, P, M, bx, by, bz);
}

// One-round launches whose tile columns get different chunk counts (balanced_map below): item k
// of the map packs the column (bx, by), the chunk index and the column's chunk count, and the tile
// body takes the chunk bounds from a map with that count.
template <int BY>
__global__ __launch_bounds__(64 * BY) __attribute__((amdgpu_waves_per_eu(4))) void velocity_dma_balanced(DevStep P,
                                                                                                    BalancedMap B) {
    __shared__ VelShared<BY> sh;
    const int b = int(blockIdx.x);
    const int logical = (b & 7) * B.per_xcd + (b >> 3);
    if (logical >= B.nblocks) return;
    const unsigned it = B.item[logical];
    TileMap M{};
    M.nbz = int(it >> 22) & 255;
    velocity_tile<BY, false>(sh, P, M, int(it & 63u), int(it >> 6) & 255, int(it >> 14) & 255);
}

template <int BY>
__global__ __launch_bounds__(64 * BY) void stress_dma_balanced(DevStep P, BalancedMap B) {
    __shared__ StrShared<BY> sh;
    const int b = int(blockIdx.x);
    const int logical = (b & 7) * B.per_xcd + (b >> 3);
    if (logical >= B.nblocks) return;
    const unsigned it = B.item[logical];
    TileMap M{};
    M.nbz = int(it >> 22) & 255;
    stress_tile<BY, false>(sh, P, M, int(it & 63u), int(it >> 6) & 255, int(it >> 14) & 255);
}

// The same half-steps for several shots in one launch: the shots' tiles are dealt out as one
// list (shot s owns logical tiles [L.first[s], L.first[s + 1])), XCD-aware like a single shot.
// L.steps[s] is shot s's step parameters for this launch (its own grid, fields and source term).
__device__ __forceinline__ bool shot_tile_of(const ShotLaunch& L, int& s, int& bx, int& by, int& bz) {
    const int b = int(blockIdx.x);
    const int logical = (b & 7) * L.per_xcd + (b >> 3);
    if (logical >= L.nblocks) return false;
    int k = 0;
    while (k + 1 < L.nshots && L.first[k + 1] <= logical) ++k;
    s = __builtin_amdgcn_readfirstlane(k);   // uniform: the shot's parameters load into SGPRs
    const TileMap& m = L.maps[s];
    const int t = logical - L.first[s];
    bx = __builtin_amdgcn_readfirstlane(t % m.nbx);
    by = __builtin_amdgcn_readfirstlane((t / m.nbx) % m.nby);
    bz = __builtin_amdgcn_readfirstlane(t / (m.nbx * m.nby));
    return true;
}

// Each shot slot has its own copy of the tile body, addressing its parameters at a fixed
// kernel-argument offset (as a single-shot launch does), so no parameter is indexed at run time.
template <int BY, int K = 0>
__device__ __forceinline__ void velocity_shot(VelShared<BY>& sh, const ShotLaunch& L, int s, int bx, int by, int bz) {
    if constexpr (K < kShotsPerLaunch) {
        if (s == K) velocity_tile<BY, false>(sh, L.steps[K], L.maps[K], bx, by, bz);
        else velocity_shot<BY, K + 1>(sh, L, s, bx, by, bz);
    }
}

template <int BY, int K = 0>
__device__ __forceinline__ void stress_shot(StrShared<BY>& sh, const ShotLaunch& L, int s, int bx, int by, int bz) {
    if constexpr (K < kShotsPerLaunch) {
        if (s == K) stress_tile<BY, false>(sh, L.steps[K], L.maps[K], bx, by, bz);
        else stress_shot<BY, K + 1>(sh, L, s, bx, by, bz);
    }
}

template <int BY>
__global__ __launch_bounds__(64 * BY) __attribute__((amdgpu_waves_per_eu(4))) void velocity_dma_shots(ShotLaunch L) {
    __shared__ VelShared<BY> sh;
    int s, bx, by, bz;
    if (!shot_tile_of(L, s, bx, by, bz)) return;
    velocity_shot<BY>(sh, L, s, bx, by, bz);
}

template <int BY>
__global__ __launch_bounds__(64 * BY) void stress_dma_shots(ShotLaunch L) {
    __shared__ StrShared<BY> sh;
    int s, bx, by, bz;
    if (!shot_tile_of(L, s, bx, by, bz)) return;
    stress_shot<BY>(sh, L, s, bx, by, bz);
}

// Work-queue half-steps: a persistent grid (every workgroup resident) pulls (shot, tile, z-chunk)
// items from per-XCD queues until they run dry.  The items of all shots form one list (shot s owns
// items [first[s], first[s + 1])), dealt to the eight XCDs in contiguous ranges of equal work
// (XCD x owns items [xcd_first[x], xcd_first[x + 1])) so that neighbouring tiles share an XCD's
// L2, as tile_of deals them; within its range an XCD's workgroups take the next item from the
// XCD's counter, so shots of uneven size balance.  counters[0..7] must be zero at launch (one set
// per launch).  Every workgroup leaves its loop as soon as its XCD's counter passes the end of
// the XCD's range, so the grid always drains.
__device__ __forceinline__ bool queue_item(const TileMap* __restrict__ maps, const int* __restrict__ first,
                                           const int* __restrict__ xcd_first, unsigned* counters, int nshots,
                                           int* slot, int& s, int& bx, int& by, int& bz) {
    const int xcd = int(blockIdx.x) & 7;
    const int lo = xcd_first[xcd];
    const int hi = xcd_first[xcd + 1];
    __syncthreads();   // every wave has finished with the previous item (its LDS and *slot)
    if (threadIdx.x == 0 && threadIdx.y == 0) *slot = lo + int(atomicAdd(&counters[xcd], 1u));
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(*slot);
    if (item >= hi) return false;
    int k = 0;
    while (k + 1 < nshots && first[k + 1] <= item) ++k;
    s = __builtin_amdgcn_readfirstlane(k);
    const TileMap& m = maps[s];
    const int t = item - first[s];
    bx = __builtin_amdgcn_readfirstlane(t % m.nbx);
    by = __builtin_amdgcn_readfirstlane((t / m.nbx) % m.nby);
    bz = __builtin_amdgcn_readfirstlane(t / (m.nbx * m.nby));
    return true;
}

// The step table read as constant memory (the kernels never write it): the tile bodies then read
// a shot's parameters with scalar loads they may repeat at will, as from the kernel arguments,
// instead of holding them in registers (read as ordinary global memory, the stress kernel needed
// 154 VGPRs and ran 15-20 % slower).
using ConstStep = const __attribute__((address_space(4))) DevStep;

template <int BY>
__global__ __launch_bounds__(64 * BY) __attribute__((amdgpu_waves_per_eu(4))) void velocity_dma_queue(
    const DevStep* __restrict__ steps, const TileMap* __restrict__ maps, const int* __restrict__ first,
    const int* __restrict__ xcd_first, unsigned* counters, int nshots) {
    __shared__ VelShared<BY> sh;
    __shared__ int slot;
    int s, bx, by, bz;
    while (queue_item(maps, first, xcd_first, counters, nshots, &slot, s, bx, by, bz))
        velocity_tile<BY, false, true>(sh, ((ConstStep*)steps)[s], maps[s], bx, by, bz);
}

template <int BY>
__global__ __launch_bounds__(64 * BY) __attribute__((amdgpu_waves_per_eu(4))) void stress_dma_queue(const DevStep* __restrict__ steps,
                                                          const TileMap* __restrict__ maps,
                                                          const int* __restrict__ first,
                                                          const int* __restrict__ xcd_first, unsigned* counters,
                                                          int nshots) {
    __shared__ StrShared<BY> sh;
    __shared__ int slot;
    int s, bx, by, bz;
    while (queue_item(maps, first, xcd_first, counters, nshots, &slot, s, bx, by, bz))
        stress_tile<BY, false, true>(sh, ((ConstStep*)steps)[s], maps[s], bx, by, bz);
}

__global__ void zero_words(unsigned* w, long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) w[i] = 0u;
}

hipError_t launch_zero_words(unsigned* w, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_words, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, w, n);
    return hipGetLastError();
}

hipError_t queue_grid(int* blocks) {
    int dev = 0, cus = 0, v = 0, st = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, velocity_dma_queue<4>, 64 * 4, 0);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&st, stress_dma_queue<4>, 64 * 4, 0);
    if (e != hipSuccess) return e;
    // Whole rounds of eight; a workgroup that finds no resident slot starts when another leaves and
    // finds the queues empty, so the count only has to cover the resident slots.
    *blocks = std::max(8, (std::max(v, st) * cus) / 8 * 8);
    return hipSuccess;
}

hipError_t launch_dma_queue(bool velocity, const DevStep* steps, const TileMap* maps, const int* first,
                            const int* xcd_first, unsigned* counters, int nshots, int blocks, hipStream_t s) {
    if (nshots < 1 || blocks < 8 || blocks % 8) return hipErrorInvalidValue;
    if (velocity)
        hipLaunchKernelGGL(velocity_dma_queue<4>, dim3(unsigned(blocks)), dim3(64, 4), 0, s, steps, maps, first,
                           xcd_first, counters, nshots);
    else
        hipLaunchKernelGGL(stress_dma_queue<4>, dim3(unsigned(blocks)), dim3(64, 4), 0, s, steps, maps, first,
                           xcd_first, counters, nshots);
    return hipGetLastError();
}

hipError_t dma_residency(int tile_y, int* blocks_per_cu, int* cus) {
    int dev = 0, v = 0, st = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    if (tile_y == 4) {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, velocity_dma<4, false>, 64 * 4, 0);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&st, stress_dma<4, false>, 64 * 4, 0);
    } else if (tile_y == 8) {
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, velocity_dma<8, false>, 64 * 8, 0);
        if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&st, stress_dma<8, false>, 64 * 8, 0);
    } else {
        return hipErrorInvalidValue;
    }
    *blocks_per_cu = std::min(v, st);
    return e;
}

// A balanced one-round map.  A tile column whose cells lie in an x or a y CPML slab costs more per
// plane than an interior one (its CPML memories are read and written, and a tile that the slab
// ends mid-way costs a whole tile), so with equal chunks the interior columns finish early and
// idle while the slab columns finish the launch.  Here each column class gets its own chunk count,
// in proportion to its cost per plane, so that every workgroup of the one round carries about the
// same work: the smallest per-workgroup work T for which the columns' chunk counts
// ceil(planes * weight / T) fit the resident slots.  Items are ordered by the depth of their
// chunk, then row, then column, and dealt to the XCDs in contiguous ranges (as tile_of deals the
// uniform map), so each XCD takes a band of depths and neighbouring tiles share its L2.
// Returns false when no balanced map fits one round (too many columns or chunks over 128 planes).
bool balanced_map(const DevGrid& g, int tile_y, int slots, BalancedMap* out) {
    constexpr double kSlabX = 0.21, kSlabY = 0.16;   // extra cost per plane, measured per class
    const int nbx = (g.gx - 8 + 63) / 64, nby = (g.gy - 8 + tile_y - 1) / tile_y, uz = g.gz - 8;
    const int cols = nbx * nby;
    slots = std::min(slots, kBalancedItems);
    const int min_chunks = (uz + kMaxChunk - 1) / kMaxChunk;
    if (cols < 1 || nbx > 64 || nby > 256 || cols * min_chunks > slots) return false;
    std::vector<double> weight(cols);
    for (int by = 0; by < nby; ++by)
        for (int bx = 0; bx < nbx; ++bx) {
            const int x0 = 4 + bx * 64, x1 = std::min(x0 + 64, g.gx - 4);
            const int y0 = 4 + by * tile_y, y1 = std::min(y0 + tile_y, g.gy - 4);
            const bool xs = g.pml > 0 && (x0 < 4 + g.pml || x1 > g.gx - 4 - g.pml);
            const bool ys = g.pml > 0 && (y0 < 4 + g.pml || y1 > g.gy - 4 - g.pml);
            weight[by * nbx + bx] = 1.0 + (xs ? kSlabX : 0.0) + (ys ? kSlabY : 0.0);
        }
    const int max_chunks = std::max(min_chunks, std::min(255, uz / 8));   // chunks of 8 planes or more
    auto chunks = [&](double work, int c) {
        return std::min(max_chunks, std::max(min_chunks, int(std::ceil(uz * weight[c] / work))));
    };
    auto total = [&](double work) {
        int n = 0;
        for (int c = 0; c < cols; ++c) n += chunks(work, c);
        return n;
    };
    double lo = 0.5, hi = 2.0 * uz;                  // total(hi) = cols * min_chunks <= slots
    for (int it = 0; it < 60; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (total(mid) <= slots) hi = mid; else lo = mid;
    }
    struct Item { double depth; int by, bx, j, n; };
    std::vector<Item> items;
    for (int c = 0; c < cols; ++c) {
        const int n = chunks(hi, c);
        for (int j = 0; j < n; ++j) items.push_back({(j + 0.5) / n, c / nbx, c % nbx, j, n});
    }
    if (int(items.size()) > slots) return false;
    std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) {
        if (a.depth != b.depth) return a.depth < b.depth;
        return a.by != b.by ? a.by < b.by : a.bx < b.bx;
    });
    out->nblocks = int(items.size());
    out->per_xcd = (out->nblocks + 7) / 8;
    for (int k = 0; k < out->nblocks; ++k) {
        const Item& i = items[k];
        out->item[k] = unsigned(i.bx) | unsigned(i.by) << 6 | unsigned(i.j) << 14 | unsigned(i.n) << 22;
    }
    return true;
}

TileMap dma_tile_map(const DevGrid& g, int tile_y, int zchunk) {
    if (zchunk <= 0) zchunk = 16;   // balanced launches (zchunk 0) step batches with the default chunks
    TileMap m;
    m.nbx = (g.gx - 8 + 63) / 64;
    m.nby = (g.gy - 8 + tile_y - 1) / tile_y;
    m.zchunk = zchunk;
    m.nbz = (g.gz - 8 + zchunk - 1) / zchunk;
    m.nblocks = m.nbx * m.nby * m.nbz;
    m.per_xcd = (m.nblocks + 7) / 8;
    return m;
}

hipError_t launch_dma_shots(bool velocity, const ShotLaunch& L, int tile_y, hipStream_t s) {
    if (L.nshots < 1 || L.nshots > kShotsPerLaunch || L.nblocks < 1) return hipErrorInvalidValue;
    const dim3 grid(unsigned(L.per_xcd) * 8);
    switch (tile_y) {
    case 4:
        if (velocity) hipLaunchKernelGGL(velocity_dma_shots<4>, grid, dim3(64, 4), 0, s, L);
        else hipLaunchKernelGGL(stress_dma_shots<4>, grid, dim3(64, 4), 0, s, L);
        break;
    case 8:
        if (velocity) hipLaunchKernelGGL(velocity_dma_shots<8>, grid, dim3(64, 8), 0, s, L);
        else hipLaunchKernelGGL(stress_dma_shots<8>, grid, dim3(64, 8), 0, s, L);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_dma(bool velocity, const DevStep& P, int tile_y, int zchunk, hipStream_t s) {
    if (zchunk > kMaxChunk) return hipErrorInvalidValue;
    const bool oop = out_of_place(P.p);
    if (zchunk == 0 && tile_y == 8 && !oop) {
        // The balanced maps of the grids this thread launches (batches alternate between shots of
        // several grids), each built once.
        struct Entry {
            int key[5];
            bool ok;
            BalancedMap map;
        };
        thread_local std::vector<Entry> cache;
        const int k[5] = {P.g.gx, P.g.gy, P.g.gz, P.g.pml, tile_y};
        const Entry* hit = nullptr;
        for (const auto& e : cache)
            if (std::equal(k, k + 5, e.key)) { hit = &e; break; }
        if (!hit) {
            int dev = 0, cus = 0, v = 0, st = 0;
            hipError_t e = hipGetDevice(&dev);
            if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
            if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&v, velocity_dma_balanced<8>, 64 * 8, 0);
            if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&st, stress_dma_balanced<8>, 64 * 8, 0);
            if (e != hipSuccess) return e;
            if (cache.size() >= 64) cache.erase(cache.begin());
            Entry fresh{};
            std::copy(k, k + 5, fresh.key);
            fresh.ok = balanced_map(P.g, tile_y, std::min(v, st) * cus, &fresh.map);
            cache.push_back(fresh);
            hit = &cache.back();
        }
        if (hit->ok) {
            const dim3 grid(unsigned(hit->map.per_xcd) * 8);
            if (velocity) hipLaunchKernelGGL(velocity_dma_balanced<8>, grid, dim3(64, 8), 0, s, P, hit->map);
            else hipLaunchKernelGGL(stress_dma_balanced<8>, grid, dim3(64, 8), 0, s, P, hit->map);
            return hipGetLastError();
        }
    }
    const TileMap m = dma_tile_map(P.g, tile_y, zchunk);
    const dim3 grid(unsigned(m.per_xcd) * 8);
    switch (tile_y) {
    case 4:
        if (oop) {
            if (velocity) hipLaunchKernelGGL((velocity_dma<4, true>), grid, dim3(64, 4), 0, s, P, m);
            else hipLaunchKernelGGL((stress_dma<4, true>), grid, dim3(64, 4), 0, s, P, m);
        } else {
            if (velocity) hipLaunchKernelGGL((velocity_dma<4, false>), grid, dim3(64, 4), 0, s, P, m);
            else hipLaunchKernelGGL((stress_dma<4, false>), grid, dim3(64, 4), 0, s, P, m);
        }
        break;
    case 8:
        if (oop) {
            if (velocity) hipLaunchKernelGGL((velocity_dma<8, true>), grid, dim3(64, 8), 0, s, P, m);
            else hipLaunchKernelGGL((stress_dma<8, true>), grid, dim3(64, 8), 0, s, P, m);
        } else {
            if (velocity) hipLaunchKernelGGL((velocity_dma<8, false>), grid, dim3(64, 8), 0, s, P, m);
            else hipLaunchKernelGGL((stress_dma<8, false>), grid, dim3(64, 8), 0, s, P, m);
        }
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace seis3d
