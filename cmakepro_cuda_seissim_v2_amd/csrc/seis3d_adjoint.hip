// seis3d_adjoint.hip - discrete adjoint of the 3D O(8) velocity-stress step
// (no reference counterpart): the FWI gradient with respect to 1/rho, lambda
// and mu, folded and chain-ruled to vp, vs and rho.
//
// For a fixed model one step is linear in (v, sigma, psi), so its adjoint is
// the transpose of the same operations in reverse order:
//  * stencils: D+^T = -D-, D-^T = -D+ (the O(8) weights are antisymmetric);
//  * CPML: psi' = b psi + a d, e = d/kappa + psi' transposes to
//    L_tot = L_psi' + L_e, L_d = L_e/kappa + a L_tot, L_psi = b L_tot;
//  * constitutive law: L_e = dt M^T L_sigma; momentum: L_e = dt h L_v.
// Each half-step is a local pass (filtered adjoint derivatives L_d into nine
// temporaries, CPML adjoint memories, gradient sums) followed by a gather
// pass applying the transposed stencils (2.5D: z-queues in registers, x/y
// halos staged in LDS, the next plane's loads in flight during the current
// one).  Gradients are summed over time per cell in their cheapest form
// (dJ/dlambda and dJ/dmu - 2 dJ/dlambda of the normal stresses, dJ/dmu_h of
// each harmonic shear modulus, dJ/dh of each face buoyancy); the harmonic-mean
// derivatives are applied once at the end.  The second normal-stress sum is
// the shear-modulus gradient at fixed P modulus, which is what dJ/dvs needs:
// each step adds only its cross terms -2 dt sum_{a != b} e_aa L_b, instead of
// two P-wave sums that would be differenced at the end.  A free surface adds the transposes of its ghost
// fills (plane-sized passes) and of its constitutive law at z = 4; reflecting
// walls add the transposes of their mirror fills.
// Built with -ffp-contract=off, as the forward kernels.
#include "seis3d_adjoint.hpp"
#include "seis3d_device.hpp"

namespace seis3d {

namespace {

__device__ __forceinline__ long long acell(const DevGrid& g, int x, int y, int z) {
    return (static_cast<long long>(z) * g.gy + y) * g.px + x;
}

template <bool plus>
__device__ __forceinline__ float aderiv(const float* __restrict__ f, long long i, long long s, float invh) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = f[i + (j - (plus ? 3 : 4)) * s];
    return stencil8(v, invh);
}

__device__ __forceinline__ int source_r2a(const DevSource& s, int x, int y, int z) {
    const int dx = x - s.x, dy = y - s.y, dz = z - s.z;
    if (dx < -1 || dx > 1 || dy < -1 || dy > 1 || dz < -1 || dz > 1) return -1;
    return dx * dx + dy * dy + dz * dz;
}

// CPML slot position of a derivative along `axis` at (x, y, z), -1 outside the slab.
__device__ __forceinline__ long long slab_index(const DevGrid& g, int axis, int slot, int q, int x, int y, int z) {
    if (q < 0) return -1;
    if (axis == 0) return psi_index_x(g, slot, q, y, z);
    if (axis == 1) return psi_index_y(g, slot, q, x, z);
    return psi_index_z(g, slot, q, x, y);
}

// Filtered derivative of a stored state: e = d/kappa + psi' (psi' after the step).
__device__ __forceinline__ float filtered(const float* __restrict__ psi, long long idx, const float* __restrict__ k,
                                          float d) {
    return k[2] * d + psi[idx];
}

// Adjoint of the CPML filter of one derivative: returns L_d, updates L_psi.
__device__ __forceinline__ float adjoint_filter(float* __restrict__ apsi, long long idx, const float* __restrict__ k,
                                                float le) {
    const float total = apsi[idx] + le;
    apsi[idx] = k[1] * total;
    return k[2] * le + k[0] * total;
}

struct Cell {
    int x, y, z;
    long long i;
    int q[3];
};

// The local passes run one thread per updated cell, a 64 x 4 tile of one plane
// per workgroup, with the tiles dealt in bricks of kBrickX x kBrickY tiles by
// kBrickZ planes: the hardware deals workgroup b to XCD b % 8, and each XCD
// takes a contiguous run of bricks, plane by plane within a brick, so that the
// planes a z-stencil reads are in that XCD's L2 when its neighbours need them.
// (Against a plain x-y-z grid this cuts their HBM reads by 35-45 %.)
constexpr int kBrickX = 4, kBrickY = 4, kBrickZ = 16;
constexpr int kBrickTiles = kBrickX * kBrickY * kBrickZ;

struct BrickGrid {
    int ntx, nty, ntz;      // tiles per axis (x: 64 cells, y: 4 cells, z: 1 plane)
    int nbx, nby, nbz;      // bricks per axis
    int per_xcd;            // launched workgroups / 8
};

__host__ __device__ inline BrickGrid brick_grid(const DevGrid& g) {
    BrickGrid b;
    b.ntx = (g.gx - 8 + 63) / 64;
    b.nty = (g.gy - 8 + 3) / 4;
    b.ntz = g.gz - 8;
    b.nbx = (b.ntx + kBrickX - 1) / kBrickX;
    b.nby = (b.nty + kBrickY - 1) / kBrickY;
    b.nbz = (b.ntz + kBrickZ - 1) / kBrickZ;
    b.per_xcd = (b.nbx * b.nby * b.nbz * kBrickTiles + 7) / 8;
    return b;
}

__device__ __forceinline__ bool updated_cell(const DevGrid& g, Cell& c) {
    const BrickGrid B = brick_grid(g);
    const int blk = int(blockIdx.x);
    const int logical = (blk & 7) * B.per_xcd + (blk >> 3);
    const int brick = logical / kBrickTiles, local = logical % kBrickTiles;
    const int tx = (brick % B.nbx) * kBrickX + local % kBrickX;
    const int ty = ((brick / B.nbx) % B.nby) * kBrickY + (local / kBrickX) % kBrickY;
    const int tz = (brick / (B.nbx * B.nby)) * kBrickZ + local / (kBrickX * kBrickY);
    c.x = tx * 64 + int(threadIdx.x) + 4;
    c.y = ty * 4 + int(threadIdx.y) + 4;
    c.z = tz + 4;
    if (tz >= B.ntz || c.x >= g.gx - 4 || c.y >= g.gy - 4 || c.z >= g.gz - 4) return false;
    c.i = acell(g, c.x, c.y, c.z);
    c.q[0] = slab_q(c.x, g.gx, g.pml, false);
    c.q[1] = slab_q(c.y, g.gy, g.pml, false);
    c.q[2] = slab_q(c.z, g.gz, g.pml, g.free_surface != 0);
    return true;
}

__device__ __forceinline__ const float* coef_row(const AdjStep& P, int axis, int half, const Cell& c) {
    const int n = axis == 0 ? P.g.gx : axis == 1 ? P.g.gy : P.g.gz;
    const int coord = axis == 0 ? c.x : axis == 1 ? c.y : c.z;
    return P.coef[axis] + (half * n + coord) * 3;
}

// ---------------------------------------------------------------------------
// 2.5D plane march of the gather passes
// ---------------------------------------------------------------------------
// A 64 x 4 workgroup owns an x-y tile of the updated cells and marches a z-chunk.
// A gather differentiates, for each component v, one field along x, one along y and
// one along z (its "triples").  The z-fields live in 9-deep register queues along the
// thread's own column, so each plane is read once instead of eight times; each plane's
// x- and y-fields are staged in LDS with their 4-cell halos; the next plane's loads
// are in flight while the current plane computes.  The local passes stay one thread
// per cell: their many per-cell loads (CPML memories, gradient sums, materials) hide
// behind 8 waves per SIMD, and marched they ran at half that speed.
constexpr int kMarchChunk = 12;

struct Triples {
    const float* x[3];
    const float* y[3];
    const float* z[3];
};

struct MarchTile {
    float sx[3][4][72];    // x-fields on rows y0..y0+3, columns x0-4..x0+67
    float sy[3][12][64];   // y-fields on rows y0-4..y0+7, columns x0..x0+63
};

// One thread's share of a plane's staged fields, held in registers between the load
// and the LDS store.
struct MarchStage {
    float x[3][2];   // columns x0-4+tx and x0+60+tx (the latter for tx < 8)
    float y[3][3];   // rows y0-4+ty, y0+ty, y0+4+ty
};

__device__ __forceinline__ void load_plane(const DevGrid& g, const Triples& F, MarchStage& S, int x0, int y0,
                                           long long plane) {
    const int tx = int(threadIdx.x), ty = int(threadIdx.y);
    const int yr = y0 + ty;
    const long long row = plane + static_cast<long long>(yr) * g.px;
    const int xa = x0 - 4 + tx, xb = x0 + 60 + tx, xc = x0 + tx;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        S.x[v][0] = yr < g.gy && xa < g.gx ? F.x[v][row + xa] : 0.0f;
        S.x[v][1] = tx < 8 && yr < g.gy && xb < g.gx ? F.x[v][row + xb] : 0.0f;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int yc = y0 - 4 + ty + 4 * k;
            S.y[v][k] = yc < g.gy && xc < g.gx ? F.y[v][plane + static_cast<long long>(yc) * g.px + xc] : 0.0f;
        }
    }
}

__device__ __forceinline__ void store_plane(MarchTile& T, const MarchStage& S) {
    const int tx = int(threadIdx.x), ty = int(threadIdx.y);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        T.sx[v][ty][tx] = S.x[v][0];
        if (tx < 8) T.sx[v][ty][64 + tx] = S.x[v][1];
#pragma unroll
        for (int k = 0; k < 3; ++k) T.sy[v][ty + 4 * k][tx] = S.y[v][k];
    }
}

// Derivatives of the staged plane at the thread's cell.
struct PlaneStencils {
    const MarchTile& T;
    const float (&q)[3][9];   // z-fields at z-4 .. z+4
    float invh;
    // Triple field (v, axis a): D+ (taps -3..4) when plus, else D- (taps -4..3).
    __device__ __forceinline__ float operator()(int v, int a, bool plus) const {
        const int tx = int(threadIdx.x), ty = int(threadIdx.y);
        const int off = plus ? 3 : 4;
        float s8[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (a == 0) s8[j] = T.sx[v][ty][tx + 4 + j - off];
            else if (a == 1) s8[j] = T.sy[v][ty + 4 + j - off][tx];
            else s8[j] = q[v][j + 4 - off];
        }
        return stencil8(s8, invh);
    }
};

// Runs body(cell, stencils) at every updated cell of the workgroup's tile and z-chunk.
template <class Body>
__device__ __forceinline__ void march(const AdjStep& P, const TileMap& M, const Triples& F, MarchTile& T,
                                      Body&& body) {
    int bx, by, bz;
    if (!tile_of(M, bx, by, bz)) return;
    const DevGrid& g = P.g;
    const int x0 = 4 + bx * 64, y0 = 4 + by * 4;
    Cell c;
    c.x = x0 + int(threadIdx.x);
    c.y = y0 + int(threadIdx.y);
    const bool own = c.x < g.gx - 4 && c.y < g.gy - 4;
    c.q[0] = slab_q(c.x, g.gx, g.pml, false);
    c.q[1] = slab_q(c.y, g.gy, g.pml, false);
    int z0, z1;
    chunk_of(M, g.gz, bz, z0, z1);
    const long long col = static_cast<long long>(c.y) * g.px + c.x;
    float q[3][9], lead[3];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
#pragma unroll
        for (int j = 0; j < 8; ++j) q[v][j + 1] = own ? F.z[v][col + static_cast<long long>(z0 - 4 + j) * g.plane] : 0.0f;
        lead[v] = own ? F.z[v][col + static_cast<long long>(z0 + 4) * g.plane] : 0.0f;
    }
    MarchStage S;
    load_plane(g, F, S, x0, y0, static_cast<long long>(z0) * g.plane);
    for (int z = z0; z < z1; ++z) {
        const long long plane = static_cast<long long>(z) * g.plane;
        c.z = z;
        c.i = plane + col;
        c.q[2] = slab_q(z, g.gz, g.pml, g.free_surface != 0);
#pragma unroll
        for (int v = 0; v < 3; ++v) shift(q[v], lead[v]);
        store_plane(T, S);
        if (z + 1 < z1) {
#pragma unroll
            for (int v = 0; v < 3; ++v) lead[v] = own ? F.z[v][c.i + 5 * g.plane] : 0.0f;
            load_plane(g, F, S, x0, y0, plane + g.plane);
        }
        __syncthreads();
        if (own) body(c, PlaneStencils{T, q, P.invh});
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Stress half-step adjoint
// ---------------------------------------------------------------------------

// Local pass: strain rates of the stored velocities, gradient sums, and the
// filtered adjoint derivatives L_d[v][a] = adjoint CPML of dt M^T L_sigma.
__global__ __launch_bounds__(256) void adjoint_stress_local(AdjStep P) {
    Cell c;
    if (!updated_cell(P.g, c)) return;
    const long long st[3] = {1, P.g.px, P.g.plane};
    float e[3][3];
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float d = a == v ? aderiv<false>(P.fv[v], c.i, st[a], P.invh) : aderiv<true>(P.fv[v], c.i, st[a], P.invh);
            const long long k = slab_index(P.g, a, 3 + v, c.q[a], c.x, c.y, c.z);
            if (k >= 0) d = filtered(P.psi[a], k, coef_row(P, a, a != v ? 1 : 0, c), d);
            e[v][a] = d;
        }
    const float lam = P.lambda[c.i], m = P.mu[c.i];
    float Ls[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) Ls[k] = P.a[3 + k][c.i];
    const float dt = P.dt;
    const float* r = P.inv_mu;
    const long long i = c.i;
    const float mxy = shear_from_inverse(r[i], r[i + st[0]], r[i + st[1]], r[i + st[0] + st[1]]);
    const float myz = shear_from_inverse(r[i], r[i + st[1]], r[i + st[2]], r[i + st[1] + st[2]]);
    const float mxz = shear_from_inverse(r[i], r[i + st[0]], r[i + st[2]], r[i + st[0] + st[2]]);
    float le[3][3];
    if (P.g.free_surface && c.z == 4) {
        // Surface law: d_zz = -r (e_xx + e_yy) turns lambda * trace into
        // l' (e_xx + e_yy), l' = 2 lambda mu / (lambda + 2 mu); szz is set to 0,
        // so its adjoint clears.
        const float den = lam + 2 * m, lp = 2 * lam * m / den;
        const float sumL = Ls[0] + Ls[1];
        le[0][0] = dt * (lp * sumL + 2 * m * Ls[0]);
        le[1][1] = dt * (lp * sumL + 2 * m * Ls[1]);
        le[2][2] = 0.0f;
        const float glp = dt * (e[0][0] + e[1][1]) * sumL;
        P.g_lam[i] += glp * 4 * m * m / (den * den);
        // dJ/dmu - 2 dJ/dlambda: 2 lambda^2 - 8 mu^2 = 2 (lambda - 2 mu)(lambda + 2 mu).
        P.g_dev[i] += glp * 2 * (lam - 2 * m) / den + dt * 2 * (e[0][0] * Ls[0] + e[1][1] * Ls[1]);
        P.a[5][i] = 0.0f;
    } else {
        const float sumL = Ls[0] + Ls[1] + Ls[2];
#pragma unroll
        for (int v = 0; v < 3; ++v) le[v][v] = dt * (lam * sumL + 2 * m * Ls[v]);
        const float tr = e[0][0] + e[1][1] + e[2][2];
        P.g_lam[i] += dt * tr * sumL;
        // dJ/dmu - 2 dJ/dlambda = 2 dt (sum_a e_aa L_a - tr sumL) = -2 dt sum_{a != b} e_aa L_b.
        P.g_dev[i] -= dt * 2 * (e[0][0] * (Ls[1] + Ls[2]) + e[1][1] * (Ls[0] + Ls[2]) + e[2][2] * (Ls[0] + Ls[1]));
    }
    le[0][1] = le[1][0] = dt * mxy * Ls[3];
    le[1][2] = le[2][1] = dt * myz * Ls[4];
    le[0][2] = le[2][0] = dt * mxz * Ls[5];
    P.g_shear[0][i] += dt * (e[0][1] + e[1][0]) * Ls[3];
    P.g_shear[1][i] += dt * (e[1][2] + e[2][1]) * Ls[4];
    P.g_shear[2][i] += dt * (e[0][2] + e[2][0]) * Ls[5];
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float ld = le[v][a];
            const long long k = slab_index(P.g, a, 3 + v, c.q[a], c.x, c.y, c.z);
            if (k >= 0) ld = adjoint_filter(P.apsi[a], k, coef_row(P, a, a != v ? 1 : 0, c), ld);
            P.t[3 * v + a][i] = ld;
        }
}

// Gather pass: L_v -= D^T-stencils of L_d (D- for a == v was used forward, so D+ here).
__global__ __launch_bounds__(256) void adjoint_stress_gather(AdjStep P, TileMap M) {
    __shared__ MarchTile T;
    const Triples F{{P.t[0], P.t[3], P.t[6]}, {P.t[1], P.t[4], P.t[7]}, {P.t[2], P.t[5], P.t[8]}};
    march(P, M, F, T, [&](const Cell& c, const PlaneStencils& D) {
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            float acc = P.a[v][c.i];
#pragma unroll
            for (int a = 0; a < 3; ++a) acc -= D(v, a, a == v);
            P.a[v][c.i] = acc;
        }
    });
}

// ---------------------------------------------------------------------------
// Velocity half-step adjoint
// ---------------------------------------------------------------------------

// Local pass: stress divergence of the stored stresses (and the force source) for the
// buoyancy gradients, and the filtered adjoint derivatives L_d[v][a] of dt h L_v.
__global__ __launch_bounds__(256) void adjoint_velocity_local(AdjStep P) {
    Cell c;
    if (!updated_cell(P.g, c)) return;
    const long long st[3] = {1, P.g.px, P.g.plane};
    const int sf[3][3] = {{0, 3, 5}, {3, 1, 4}, {5, 4, 2}};   // stress index - 3
    const int r2 = P.src.active ? source_r2a(P.src, c.x, c.y, c.z) : -1;
    const float dt = P.dt;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        float E = 0.0f;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float* fld = P.fs[sf[v][a]];
            float d = a == v ? aderiv<true>(fld, c.i, st[a], P.invh) : aderiv<false>(fld, c.i, st[a], P.invh);
            const long long k = slab_index(P.g, a, v, c.q[a], c.x, c.y, c.z);
            if (k >= 0) d = filtered(P.psi[a], k, coef_row(P, a, a == v ? 1 : 0, c), d);
            E += d;
        }
        const float lv = P.a[v][c.i];
        const float h = face_buoyancy(P.inv_rho[c.i], P.inv_rho[c.i + st[v]]);
        float dvdh = dt * E;
        if (r2 >= 0) dvdh += (P.src.scale * P.src.w[r2] / P.src.norm) * P.src.dir[v];
        P.g_buoy[v][c.i] += dvdh * lv;
        const float le = dt * h * lv;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float ld = le;
            const long long k = slab_index(P.g, a, v, c.q[a], c.x, c.y, c.z);
            if (k >= 0) ld = adjoint_filter(P.apsi[a], k, coef_row(P, a, a == v ? 1 : 0, c), ld);
            P.t[3 * v + a][c.i] = ld;
        }
    }
}

// Gather pass: L_sigma -= transposed stencils of L_d, in the (v, a) order of the forward sum.
__global__ __launch_bounds__(256) void adjoint_velocity_gather(AdjStep P, TileMap M) {
    __shared__ MarchTile T;
    const Triples F{{P.t[0], P.t[3], P.t[6]}, {P.t[1], P.t[4], P.t[7]}, {P.t[2], P.t[5], P.t[8]}};
    march(P, M, F, T, [&](const Cell& c, const PlaneStencils& D) {
        const int sf[3][3] = {{0, 3, 5}, {3, 1, 4}, {5, 4, 2}};
        float s[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) s[k] = P.a[3 + k][c.i];
#pragma unroll
        for (int v = 0; v < 3; ++v)
#pragma unroll
            for (int a = 0; a < 3; ++a) s[sf[v][a]] -= D(v, a, a != v);   // D- for a == v, D+ otherwise
#pragma unroll
        for (int k = 0; k < 6; ++k) P.a[3 + k][c.i] = s[k];
    });
}

// ---------------------------------------------------------------------------
// Free surface: transposes of stressGhostKernel3D and velocityGhostKernel3D
// (cuda/Cuda_Seis3D.cu:203-219, 249-272), one thread per surface column
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool surface_column(const DevGrid& g, int& x, int& y) {
    x = int(blockIdx.x * blockDim.x + threadIdx.x) + 4;
    y = int(blockIdx.y * blockDim.y + threadIdx.y) + 4;
    return x < g.gx - 4 && y < g.gy - 4;
}

// Each stress ghost adjoint moves onto the cell it mirrors, with the ghost's sign,
// and clears; then the adjoint of the surface szz (set to zero) clears.
__global__ void adjoint_stress_ghost(AdjStep P) {
    int x, y;
    if (!surface_column(P.g, x, y)) return;
    const DevGrid& g = P.g;
#pragma unroll
    for (int depth = 1; depth <= 4; ++depth) {
        const long long ghost = acell(g, x, y, 4 - depth), mirror = acell(g, x, y, 4 + depth);
        const long long shear_mirror = acell(g, x, y, 3 + depth);
        P.a[5][mirror] -= P.a[5][ghost];
        P.a[8][shear_mirror] -= P.a[8][ghost];
        P.a[7][shear_mirror] -= P.a[7][ghost];
        P.a[3][mirror] += P.a[3][ghost];
        P.a[4][mirror] += P.a[4][ghost];
        P.a[6][mirror] += P.a[6][ghost];
#pragma unroll
        for (int k = 3; k < 9; ++k) P.a[k][ghost] = 0.0f;
    }
    P.a[5][acell(g, x, y, 4)] = 0.0f;
}

// The z-stencils of the rows below the surface read the ghost planes.  Their
// transposes gather the filtered adjoint derivatives (zero above z = 4) into the
// ghost adjoints: L_v after the stress half-step (D- on v's own axis), L_sigma
// after the velocity half-step (D+ on v's own axis).
template <bool VELOCITY>
__global__ void adjoint_ghost_gather(AdjStep P) {
    int x, y;
    if (!surface_column(P.g, x, y)) return;
    const DevGrid& g = P.g;
    const int zfield[3] = {5, 4, 2};   // stress index - 3 of sxz, syz, szz
#pragma unroll
    for (int zg = 0; zg < 4; ++zg) {
        const long long i = acell(g, x, y, zg);
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            const bool plus = VELOCITY ? v != 2 : v == 2;
            const int off = plus ? 3 : 4;
            float s8[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int z = zg + j - off;
                s8[j] = z >= 4 ? P.t[3 * v + 2][acell(g, x, y, z)] : 0.0f;
            }
            float* dst = VELOCITY ? P.a[3 + zfield[v]] : P.a[v];
            dst[i] -= stencil8(s8, P.invh);
        }
    }
}

// First half of the velocity-ghost transpose, per surface column: the ghost adjoints
// move onto the cells they copy and clear; the adjoints of the three surface slopes
// dzVx, dzVy, dzVz (times 1/h, dzVz's times -r) go to the scratch planes; and
// r = lambda / (lambda + 2 mu) takes its gradient dJ/dr = -(ex + ey) L_dzVz.
__global__ void adjoint_velocity_ghost_fold(AdjStep P) {
    int x, y;
    if (!surface_column(P.g, x, y)) return;
    const DevGrid& g = P.g;
    const float h = P.spacing;
    float l0 = 0.0f, l1 = 0.0f, l2 = 0.0f;
#pragma unroll
    for (int depth = 1; depth <= 4; ++depth) {
        const long long ghost = acell(g, x, y, 4 - depth), mirror = acell(g, x, y, 4 + depth);
        const float g0 = P.a[0][ghost], g1 = P.a[1][ghost], g2 = P.a[2][ghost];
        P.a[0][mirror] += g0;
        P.a[1][mirror] += g1;
        P.a[2][acell(g, x, y, 3 + depth)] += g2;
        l0 -= float(2 * depth) * h * g0;
        l1 -= float(2 * depth) * h * g1;
        l2 -= float(2 * depth - 1) * h * g2;
        P.a[0][ghost] = 0.0f; P.a[1][ghost] = 0.0f; P.a[2][ghost] = 0.0f;
    }
    const long long i = acell(g, x, y, 4);
    const float lam = P.lambda[i], m = P.mu[i], den = lam + 2 * m;
    const float ex = aderiv<false>(P.fv[0], i, 1, P.invh);
    const float ey = aderiv<false>(P.fv[1], i, g.px, P.invh);
    const float gr = -(ex + ey) * l2;
    P.g_lam[i] += gr * 2 * m / (den * den);
    // dr/dmu at fixed lambda + 2 mu is -2 / (lambda + 2 mu).
    P.g_dev[i] -= gr * 2 / den;
    const long long col = static_cast<long long>(y) * g.px + x;
    P.surf[0][col] = P.invh * l0;
    P.surf[1][col] = P.invh * l1;
    P.surf[2][col] = -(lam / den) * P.invh * l2;
}

// Second half: the transposed surface stencils of the scratch planes (zero outside
// the surface columns).  dzVx = -1.5 D+x vz(4) + 0.5 D+x vz(5), dzVy likewise along
// y, and dzVz = -r (D-x vx(4) + D-y vy(4)).
__global__ void adjoint_velocity_ghost_spread(AdjStep P) {
    int x, y;
    if (!surface_column(P.g, x, y)) return;
    const DevGrid& g = P.g;
    const long long col = static_cast<long long>(y) * g.px + x, row = g.px;
    const float* s0 = P.surf[0];
    const float* s1 = P.surf[1];
    const float* s2 = P.surf[2];
    const float w[4] = {kW0, kW1, kW2, kW3};
    float sx = 0.0f, sy = 0.0f, tx = 0.0f, ty = 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        sx += w[r] * (s0[col - r - 1] - s0[col + r]);
        sy += w[r] * (s1[col - (r + 1) * row] - s1[col + r * row]);
        tx += w[r] * (s2[col - r] - s2[col + r + 1]);
        ty += w[r] * (s2[col - r * row] - s2[col + (r + 1) * row]);
    }
    const long long i = acell(g, x, y, 4);
    P.a[2][i] += -1.5f * sx + -1.5f * sy;
    P.a[2][i + g.plane] += 0.5f * sx + 0.5f * sy;
    P.a[0][i] += tx;
    P.a[1][i] += ty;
}

// ---------------------------------------------------------------------------
// Reflecting walls: transpose of reflectingGhostKernel3D (cuda/Cuda_Seis3D.cu:152-175)
// ---------------------------------------------------------------------------

// One axis per launch, axes in reverse order of the forward's.  On each line along
// the axis the zeroed cells' adjoints clear and each ghost adjoint moves onto its
// source, with the ghost's sign, and clears.
__global__ void adjoint_reflecting_walls(AdjStep P, int axis, int velocities) {
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
        float* L = P.a[field];
        if (half) { p[axis] = high; L[acell(g, p[0], p[1], p[2])] = 0.0f; }
        for (int side = 1; side >= 0; --side) for (int depth = 4; depth >= 1; --depth) {
            p[axis] = side ? high + depth : 4 - depth;
            const long long destination = acell(g, p[0], p[1], p[2]);
            if (!(half && !side && depth == 1)) {
                p[axis] = side ? high + 1 - depth - int(half) : 4 + depth - 1 - int(half);
                const long long source = acell(g, p[0], p[1], p[2]);
                L[source] += half ? -L[destination] : L[destination];
            }
            L[destination] = 0.0f;
        }
    }
}

// With reflecting walls every face's ghost cells are read by the stencils of the
// cells next to it.  One launch per axis: threads over the face's updated (u, v)
// cells, both sides and all four depths, gathering the normal-axis stencil of the
// temporaries (zero outside the updated cells) into the ghost adjoints.
template <bool VELOCITY>
__global__ void adjoint_face_gather(AdjStep P, int axis) {
    const DevGrid& g = P.g;
    const int lengths[3] = {g.gx, g.gy, g.gz};
    const int u_axis = (axis + 1) % 3, v_axis = (axis + 2) % 3;
    const int u = int(blockIdx.x * blockDim.x + threadIdx.x) + 4;
    const int v = int(blockIdx.y * blockDim.y + threadIdx.y) + 4;
    if (u >= lengths[u_axis] - 4 || v >= lengths[v_axis] - 4) return;
    const int n = lengths[axis];
    const long long stride = axis == 0 ? 1 : axis == 1 ? g.px : g.plane;
    const int zfield[3][3] = {{0, 3, 5}, {3, 1, 4}, {5, 4, 2}};   // stress index - 3
    int p[3] = {0, 0, 0};
    p[u_axis] = u; p[v_axis] = v;
    for (int side = 0; side < 2; ++side)
        for (int depth = 0; depth < 4; ++depth) {
            const int coord = side ? n - 4 + depth : depth;
            p[axis] = coord;
            const long long i = acell(g, p[0], p[1], p[2]);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const bool plus = VELOCITY ? axis != c : axis == c;
                const int off = plus ? 3 : 4;
                const float* t = P.t[3 * c + axis];
                float s8[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int k = coord + j - off;
                    s8[j] = k >= 4 && k < n - 4 ? t[i + (j - off) * stride] : 0.0f;
                }
                float* dst = VELOCITY ? P.a[3 + zfield[c][axis]] : P.a[c];
                dst[i] -= stencil8(s8, P.invh);
            }
        }
}

// R^T r: the receiver collocation 0.5 (v[i] + v[i - e_v]) transposed.
__global__ void adjoint_inject(AdjStep P, const int* __restrict__ rec, int count, const float* __restrict__ res) {
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= 3 * count) return;
    const int r = k / 3, v = k % 3;
    const long long i = acell(P.g, rec[3 * r], rec[3 * r + 1], rec[3 * r + 2]);
    const long long st = v == 0 ? 1 : v == 1 ? P.g.px : P.g.plane;
    const float h = 0.5f * res[3 * r + v];
    atomicAdd(P.a[v] + i, h);
    atomicAdd(P.a[v] + i - st, h);
}

__global__ void residual_kernel(const float* __restrict__ g, const float* __restrict__ o, float* __restrict__ r,
                                long long n) {
    const long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) r[k] = g[k] - o[k];
}

// ---------------------------------------------------------------------------
// Finalisation
// ---------------------------------------------------------------------------

__device__ __forceinline__ bool interior(const DevGrid& g, int x, int y, int z) {
    return x >= 4 && x < g.gx - 4 && y >= 4 && y < g.gy - 4 && z >= 4 && z < g.gz - 4;
}

// Per padded cell k: dJ/d(1/rho)[k] from the two faces that use rho[k] on each
// axis (d(2ab/(a+b))/da = 2 b^2/(a+b)^2), and dJ/dmu[k] including the four
// harmonic shear moduli of each plane that use mu[k]
// (d(4/sum 1/mu_j)/dmu_k = mu_h^2 / (4 mu_k^2)).
__global__ void adjoint_cells(AdjStep P, float* __restrict__ out_ib, float* __restrict__ out_dev) {
    const DevGrid& g = P.g;
    const int x = int(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = int(blockIdx.y);
    const int z = int(blockIdx.z);
    if (x >= g.gx) return;
    const long long k = acell(g, x, y, z);
    const long long st[3] = {1, g.px, g.plane};
    const int p[3] = {x, y, z};
    float gib = 0.0f;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        if (interior(g, x, y, z)) {
            const float aa = P.inv_rho[k], bb = P.inv_rho[k + st[v]];
            const float den = (aa + bb) * (aa + bb);
            gib += P.g_buoy[v][k] * 2 * bb * bb / den;
        }
        int q[3] = {p[0], p[1], p[2]};
        q[v] -= 1;
        if (interior(g, q[0], q[1], q[2])) {
            const long long km = k - st[v];
            const float aa = P.inv_rho[km], bb = P.inv_rho[k];
            const float den = (aa + bb) * (aa + bb);
            gib += P.g_buoy[v][km] * 2 * aa * aa / den;
        }
    }
    float gdev = interior(g, x, y, z) ? P.g_dev[k] : 0.0f;
    const int pairs[3][2] = {{0, 1}, {1, 2}, {0, 2}};
    const float* r = P.inv_mu;
    const float mk = P.mu[k];
#pragma unroll
    for (int pp = 0; pp < 3; ++pp) {
        const int A = pairs[pp][0], B = pairs[pp][1];
#pragma unroll
        for (int o = 0; o < 4; ++o) {
            int q[3] = {p[0], p[1], p[2]};
            q[A] -= o & 1;
            q[B] -= o >> 1;
            if (!interior(g, q[0], q[1], q[2])) continue;
            const long long i = acell(g, q[0], q[1], q[2]);
            const float G = P.g_shear[pp][i];
            const float mh = shear_from_inverse(r[i], r[i + st[A]], r[i + st[B]], r[i + st[A] + st[B]]);
            if (mh != 0.0f) gdev += G * mh * mh / (4 * mk * mk);
        }
    }
    out_ib[k] = gib;
    out_dev[k] = gdev;
}

// Folds the clamp-extended padding onto each physical cell (in padded raster
// order) and applies the chain rule.  The sums are taken in the moduli
// M = lambda + 2 mu (g_lam = dJ/dM at fixed mu) and mu at fixed M (g_dev), with
// M = rho vp^2, mu = rho vs^2 and 1/rho, so no difference of two large sums is
// ever formed: dJ/dvs = 2 rho vs g_dev, dJ/dvp = 2 rho vp g_lam and
// dJ/drho = vp^2 g_lam + vs^2 g_dev - g_ib / rho^2.
__global__ void adjoint_fold(FoldParams F) {
    const int x = int(blockIdx.x * blockDim.x + threadIdx.x);
    const int y = int(blockIdx.y);
    const int z = int(blockIdx.z);
    if (x >= F.nx) return;
    const DevGrid& g = F.g;
    const int x0 = x == 0 ? 0 : x + F.ox, x1 = x == F.nx - 1 ? g.gx - 1 : x + F.ox;
    const int y0 = y == 0 ? 0 : y + F.oy, y1 = y == F.ny - 1 ? g.gy - 1 : y + F.oy;
    const int z0 = z == 0 ? 0 : z + F.oz, z1 = z == F.nz - 1 ? g.gz - 1 : z + F.oz;
    float gib = 0.0f, glam = 0.0f, gdev = 0.0f;
    for (int zz = z0; zz <= z1; ++zz)
        for (int yy = y0; yy <= y1; ++yy)
            for (int xx = x0; xx <= x1; ++xx) {
                const long long i = acell(g, xx, yy, zz);
                gib += F.g_ib[i];
                glam += F.g_lam[i];
                gdev += F.g_dev[i];
            }
    const long long i = acell(g, x + F.ox, y + F.oy, z + F.oz);
    const float rho = 1.0f / F.inv_rho[i];
    const float vs2 = F.mu[i] / rho, vp2 = (F.lambda[i] + 2 * F.mu[i]) / rho;
    const float vp = sqrtf(vp2), vs = sqrtf(vs2);
    const long long j = (static_cast<long long>(z) * F.ny + y) * F.nx + x;
    F.g_vp[j] = glam * 2 * rho * vp;
    F.g_vs[j] = gdev * 2 * rho * vs;
    F.g_rho[j] = glam * vp2 + gdev * vs2 - gib / (rho * rho);
}

// ---------------------------------------------------------------------------
// Array utilities (fills instead of hipMemsetAsync, so they are graph-safe)
// ---------------------------------------------------------------------------

template <int OP>
__global__ void array_kernel(ArrayBatch b) {
    const int k = int(blockIdx.y);
    const long long n = b.n[k];
    float* __restrict__ d = b.dst[k];
    const float* __restrict__ s = b.src[k];
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    const long long n4 = n / 4;
    float4* d4 = reinterpret_cast<float4*>(d);
    const float4* s4 = reinterpret_cast<const float4*>(s);
    for (long long j = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; j < n4; j += stride) {
        if (OP == 0) d4[j] = s4[j];
        if (OP == 1) d4[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (OP == 2) {
            float4 a = d4[j];
            const float4 c = s4[j];
            a.x += c.x; a.y += c.y; a.z += c.z; a.w += c.w;
            d4[j] = a;
        }
    }
    for (long long j = 4 * n4 + static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; j < n; j += stride) {
        if (OP == 0) d[j] = s[j];
        if (OP == 1) d[j] = 0.0f;
        if (OP == 2) d[j] += s[j];
    }
}

template <int OP>
hipError_t launch_arrays(const ArrayBatch& b, hipStream_t s) {
    if (b.count <= 0) return hipSuccess;
    if (b.count > 16) return hipErrorInvalidValue;
    for (int k = 0; k < b.count; ++k)
        if ((reinterpret_cast<unsigned long long>(b.dst[k]) & 15) != 0 ||
            (OP != 1 && (reinterpret_cast<unsigned long long>(b.src[k]) & 15) != 0))
            return hipErrorInvalidValue;
    long long most = 0;
    for (int k = 0; k < b.count; ++k) most = most > b.n[k] ? most : b.n[k];
    if (most == 0) return hipSuccess;
    const long long blocks = (most / 4 + 255) / 256;
    const unsigned gx = unsigned(blocks < 2048 ? (blocks > 0 ? blocks : 1) : 2048);
    hipLaunchKernelGGL(array_kernel<OP>, dim3(gx, unsigned(b.count)), dim3(256), 0, s, b);
    return hipGetLastError();
}

// 64 x 4 tiles of the updated cells and z-chunks of at most kMarchChunk planes, dealt to
// the XCDs by tile_of.
TileMap march_tiles(const DevGrid& g) {
    TileMap m;
    m.nbx = (g.gx - 8 + 63) / 64;
    m.nby = (g.gy - 8 + 3) / 4;
    m.zchunk = kMarchChunk;
    m.nbz = (g.gz - 8 + kMarchChunk - 1) / kMarchChunk;
    m.nblocks = m.nbx * m.nby * m.nbz;
    m.per_xcd = (m.nblocks + 7) / 8;
    return m;
}

// Transposed walls of one field set: axes 2, 1, 0 (the forward runs 0, 1, 2).
void launch_walls_adjoint(const AdjStep& P, bool velocities, hipStream_t s) {
    const int lengths[3] = {P.g.gx, P.g.gy, P.g.gz};
    const dim3 block(16, 16);
    for (int axis = 2; axis >= 0; --axis) {
        const dim3 grid((lengths[(axis + 1) % 3] + 15) / 16, (lengths[(axis + 2) % 3] + 15) / 16);
        hipLaunchKernelGGL(adjoint_reflecting_walls, grid, block, 0, s, P, axis, velocities ? 1 : 0);
    }
}

template <bool VELOCITY>
void launch_face_gathers(const AdjStep& P, hipStream_t s) {
    const int lengths[3] = {P.g.gx, P.g.gy, P.g.gz};
    const dim3 block(16, 16);
    for (int axis = 0; axis < 3; ++axis) {
        const dim3 grid((lengths[(axis + 1) % 3] - 8 + 15) / 16, (lengths[(axis + 2) % 3] - 8 + 15) / 16);
        hipLaunchKernelGGL(adjoint_face_gather<VELOCITY>, grid, block, 0, s, P, axis);
    }
}

// One thread per updated cell, in bricks (updated_cell): 8 * per_xcd workgroups of 64 x 4.
dim3 updated_grid(const DevGrid& g) { return dim3(unsigned(brick_grid(g).per_xcd * 8)); }

}  // namespace

hipError_t launch_adjoint_stress(const AdjStep& P, hipStream_t s) {
    const TileMap m = march_tiles(P.g);
    const dim3 grid(unsigned(m.per_xcd * 8)), block(64, 4);
    const dim3 surface(unsigned((P.g.gx - 8 + 63) / 64), unsigned((P.g.gy - 8 + 3) / 4));
    const bool free = P.g.free_surface != 0;
    if (free) hipLaunchKernelGGL(adjoint_stress_ghost, surface, block, 0, s, P);
    if (P.reflecting) launch_walls_adjoint(P, false, s);
    hipLaunchKernelGGL(adjoint_stress_local, updated_grid(P.g), block, 0, s, P);
    hipLaunchKernelGGL(adjoint_stress_gather, grid, block, 0, s, P, m);
    if (free) {
        hipLaunchKernelGGL(adjoint_ghost_gather<false>, surface, block, 0, s, P);
        hipLaunchKernelGGL(adjoint_velocity_ghost_fold, surface, block, 0, s, P);
        hipLaunchKernelGGL(adjoint_velocity_ghost_spread, surface, block, 0, s, P);
    }
    if (P.reflecting) {
        launch_face_gathers<false>(P, s);
        launch_walls_adjoint(P, true, s);
    }
    return hipGetLastError();
}

hipError_t launch_adjoint_velocity(const AdjStep& P, hipStream_t s) {
    const TileMap m = march_tiles(P.g);
    const dim3 grid(unsigned(m.per_xcd * 8)), block(64, 4);
    const dim3 surface(unsigned((P.g.gx - 8 + 63) / 64), unsigned((P.g.gy - 8 + 3) / 4));
    hipLaunchKernelGGL(adjoint_velocity_local, updated_grid(P.g), block, 0, s, P);
    hipLaunchKernelGGL(adjoint_velocity_gather, grid, block, 0, s, P, m);
    if (P.g.free_surface) {
        hipLaunchKernelGGL(adjoint_ghost_gather<true>, surface, block, 0, s, P);
        // The step opens with stress ghosts as well.
        hipLaunchKernelGGL(adjoint_stress_ghost, surface, block, 0, s, P);
    }
    if (P.reflecting) {
        launch_face_gathers<true>(P, s);
        // The step opens with stress walls as well.
        launch_walls_adjoint(P, false, s);
    }
    return hipGetLastError();
}

hipError_t launch_adjoint_inject(const AdjStep& P, const int* rec, int count, const float* residual, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(adjoint_inject, dim3((3 * count + 255) / 256), dim3(256), 0, s, P, rec, count, residual);
    return hipGetLastError();
}

hipError_t launch_residual(const float* gather, const float* observed, float* residual, long long n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(residual_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, s, gather, observed, residual, n);
    return hipGetLastError();
}

hipError_t launch_adjoint_cells(const AdjStep& P, float* g_ib, float* out_dev, hipStream_t s) {
    hipLaunchKernelGGL(adjoint_cells, dim3(unsigned((P.g.gx + 63) / 64), unsigned(P.g.gy), unsigned(P.g.gz)), dim3(64),
                       0, s, P, g_ib, out_dev);
    return hipGetLastError();
}

hipError_t launch_adjoint_fold(const FoldParams& F, hipStream_t s) {
    hipLaunchKernelGGL(adjoint_fold, dim3(unsigned((F.nx + 63) / 64), unsigned(F.ny), unsigned(F.nz)), dim3(64), 0, s, F);
    return hipGetLastError();
}

hipError_t launch_copy_arrays(const ArrayBatch& b, hipStream_t s) { return launch_arrays<0>(b, s); }
hipError_t launch_zero_arrays(const ArrayBatch& b, hipStream_t s) { return launch_arrays<1>(b, s); }
hipError_t launch_add_arrays(const ArrayBatch& b, hipStream_t s) { return launch_arrays<2>(b, s); }

}  // namespace seis3d
