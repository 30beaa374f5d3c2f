// seis3d_device.hpp - device-side data layout shared by the 3D kernels.
//
// HBM layout (all float32):
//  * Wavefields and materials: padded reference grid gx*gy*gz
//    (include/Common3D.h:96-98), x fastest, but with the row pitch `px`
//    rounded up to 64 floats so that every row starts on a 256-byte
//    boundary and a 64-lane wavefront reads whole 128-byte lines.
//  * CPML memories: one array per axis, six slots per array (slots 0-2 are
//    the velocity derivatives, 3-5 the stress derivatives, as the
//    reference's applyCpml3D slot argument, cuda/Cuda_Seis3D.cu:92-111).
//    Unlike the reference's slab layout (z,y,q / z,x,q / y,x,q), every slab
//    here keeps x fastest so that psi traffic is coalesced:
//      axis 0: [slot][z][y][q]  with q over the 2p x-slab columns
//      axis 1: [slot][z][q][x]  with q over the 2p y-slab rows, pitch px
//      axis 2: [slot][q][y][x]  with q over the (1 or 2)p z-slab planes
//  * CPML tables: [half][n][a, b, 1/kappa] per axis, as makeCoefficients3D.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace seis3d {

struct DevGrid {
    int gx, gy, gz;       // padded reference dimensions
    int px;               // row pitch in floats (multiple of 64, >= gx)
    long long plane;      // px * gy
    int pml;              // active CPML thickness (0 when reflecting)
    int free_surface;
    int tz;               // z-slab thickness: (free ? 1 : 2) * pml
    long long slab[3];    // floats per slot of each axis' CPML array
};

struct DevFields {
    float* f[9];          // vx vy vz sxx syy szz sxy syz sxz (read)
    float* inv_rho;
    float* lambda;
    float* mu;
    float* inv_mu;        // 1/mu per cell, rebuilt after every material change
    float* psi[3];        // CPML memories (read)
    const float* coef[3];
    unsigned int* flags;  // [0] non-finite flag
    // Where a half-step writes the fields it updates and its CPML memories.  Equal to
    // f / psi for the usual in-place step; a half-step reads its own cell's old value
    // before writing it, so distinct arrays give an out-of-place step with the same
    // arithmetic.  Only the LDS-DMA kernels (family 4) take distinct arrays.
    float* out[9];
    float* psi_out[3];
};

// True when a step writes somewhere other than its inputs (an out-of-place step).
inline bool out_of_place(const DevFields& p) {
    for (int k = 0; k < 9; ++k)
        if (p.out[k] != p.f[k]) return true;
    for (int a = 0; a < 3; ++a)
        if (p.psi_out[a] != p.psi[a]) return true;
    return false;
}

// One source term of sourceKernel3D (cuda/Cuda_Seis3D.cu:313-332); the
// time function is evaluated on the host in FP64 and rounded to float.
struct DevSource {
    int active;           // inject in this launch
    int x, y, z;          // padded centre
    float scale;          // (dt * amplitude) * ricker, float arithmetic
    float w[4];           // expf(-0.5 * r^2) for r^2 = 0..3
    float norm;           // 10.8387785335f
    float dir[3];
};

struct DevStep {
    DevGrid g;
    DevFields p;
    float dt;
    float invh;
    DevSource src;
};

// weight3D (cuda/Cuda_Seis3D.cu:72-75).
__device__ constexpr float kW0 = 1225.0f / 1024.0f;
__device__ constexpr float kW1 = -245.0f / 3072.0f;
__device__ constexpr float kW2 = 49.0f / 5120.0f;
__device__ constexpr float kW3 = -5.0f / 7168.0f;

// Slab-local index q of a coordinate on one axis, -1 outside the slab
// (applyCpml3D, cuda/Cuda_Seis3D.cu:95-102).
__device__ __forceinline__ int slab_q(int coord, int n, int pml, bool top_free) {
    if (!top_free && coord < 4 + pml) return coord - 4;
    if (coord >= n - 4 - pml) return (top_free ? 0 : pml) + coord - (n - 4 - pml);
    return -1;
}

template <class G>
__device__ __forceinline__ long long psi_index_x(const G& g, int slot, int q, int y, int z) {
    return ((static_cast<long long>(slot) * g.gz + z) * g.gy + y) * (2 * g.pml) + q;
}
template <class G>
__device__ __forceinline__ long long psi_index_y(const G& g, int slot, int q, int x, int z) {
    return ((static_cast<long long>(slot) * g.gz + z) * (2 * g.pml) + q) * g.px + x;
}
template <class G>
__device__ __forceinline__ long long psi_index_z(const G& g, int slot, int q, int x, int y) {
    return ((static_cast<long long>(slot) * g.tz + q) * g.gy + y) * g.px + x;
}

// CPML recurrence (cuda/Cuda_Seis3D.cu:109-110): psi = b*psi + a*d;
// return d/kappa + psi.  Built without FP contraction.
__device__ __forceinline__ float cpml_update(float* __restrict__ psi, long long idx,
                                             const float* __restrict__ k, float d) {
    const float m = k[1] * psi[idx] + k[0] * d;
    psi[idx] = m;
    return k[2] * d + m;
}

// Harmonic face average used for buoyancy (cuda/Cuda_Seis3D.cu:240-241).
__device__ __forceinline__ float face_buoyancy(float a, float b) { return 2 * a * b / (a + b); }

// shearMu3D (cuda/Cuda_Seis3D.cu:275-278).
__device__ __forceinline__ float shear_harmonic(float m0, float m1, float m2, float m3) {
    return (m0 > 0 && m1 > 0 && m2 > 0 && m3 > 0) ? 4 / (1 / m0 + 1 / m1 + 1 / m2 + 1 / m3) : 0.0f;
}

// shearMu3D from the reciprocals r_k = 1/mu_k, each the same IEEE division
// the reference performs, so the result is bitwise identical: r > 0 holds for
// every mu > 0 (r = +inf for mu = +0 or a tiny denormal, where the reference
// also returns 4/inf = 0 or fails its test and returns 0).
__device__ __forceinline__ float shear_from_inverse(float r0, float r1, float r2, float r3) {
    return (r0 > 0 && r1 > 0 && r2 > 0 && r3 > 0) ? 4 / (r0 + r1 + r2 + r3) : 0.0f;
}

// derivative3D (cuda/Cuda_Seis3D.cu:78-89) on eight samples s[0..7] ordered
// from the most negative to the most positive tap: D+ at i+1/2 uses
// f[i-3..i+4], D- at i uses f[i-4..i+3]; either way s[4+r] - s[3-r].
__device__ __forceinline__ float stencil8(const float s[8], float invh) {
    float r = 0.0f;
    r += kW0 * (s[4] - s[3]);
    r += kW1 * (s[5] - s[2]);
    r += kW2 * (s[6] - s[1]);
    r += kW3 * (s[7] - s[0]);
    return r * invh;
}

// ---- 2.5D streaming tiles -------------------------------------------------

constexpr int TX = 64;     // one wavefront across x
constexpr int HX = TX + 8; // x extent with a 4-cell halo each side

struct TileMap {
    int nbx, nby, nbz;     // tiles per axis
    int nblocks;           // nbx * nby * nbz
    int per_xcd;           // launched blocks / 8
    int zchunk;            // planes per tile
};

// XCD-aware tile assignment: the hardware deals workgroup b to XCD b % 8, so
// logical tile (b % 8) * per_xcd + b / 8 gives each XCD a contiguous run of
// x-fastest tiles (neighbours share halos in that XCD's L2).
__device__ __forceinline__ bool tile_of(const TileMap& m, int& bx, int& by, int& bz) {
    const int b = int(blockIdx.x);
    const int logical = (b & 7) * m.per_xcd + (b >> 3);
    if (logical >= m.nblocks) return false;
    bx = logical % m.nbx;
    const int t = logical / m.nbx;
    by = t % m.nby;
    bz = t / m.nby;
    return true;
}

// Planes [z0, z1) of z-chunk bz: the updated depth gz - 8 split into nbz
// chunks whose sizes differ by at most one (never more than zchunk planes).
__device__ __forceinline__ void chunk_of(const TileMap& m, int gz, int bz, int& z0, int& z1) {
    const int uz = gz - 8;
    z0 = 4 + (bz * uz) / m.nbz;
    z1 = 4 + ((bz + 1) * uz) / m.nbz;
}

template <int N>
__device__ __forceinline__ void shift(float (&q)[N], float lead) {
#pragma unroll
    for (int j = 0; j < N - 1; ++j) q[j] = q[j + 1];
    q[N - 1] = lead;
}

// A one-round launch whose tile columns have their own chunk counts (balanced_map in
// seis3d_dma.hip): item k = bx | by << 6 | chunk << 14 | chunks << 22, dealt like tile_of.
constexpr int kBalancedItems = 512;
struct BalancedMap {
    int nblocks;
    int per_xcd;           // launched blocks / 8
    unsigned item[kBalancedItems];
};

// Several shots' tiles in one launch, passed by value as the kernel argument, so that every
// shot's parameters are read from the kernel-argument segment exactly as a single shot's are
// (constant memory: its field pointers are known to address global memory).  Shot s owns
// logical tiles [first[s], first[s + 1]).
constexpr int kShotsPerLaunch = 4;
struct ShotLaunch {
    int nshots;
    int nblocks;           // first[nshots]
    int per_xcd;           // launched blocks / 8
    int first[kShotsPerLaunch + 1];
    TileMap maps[kShotsPerLaunch];
    DevStep steps[kShotsPerLaunch];
};
static_assert(sizeof(ShotLaunch) <= 4096, "kernel argument segment");

}  // namespace seis3d
