// seis3d_abc.hip - learned absorbing boundary kernels (no reference
// counterpart): frame-cell features, the BF16 MFMA MLP (inference and
// training forward), FP32 backward, parameter-gradient reductions and Adam.
//
// MFMA: v_mfma_f32_16x16x16_bf16 on 64-lane wavefronts.  Lane l supplies
// A[row l%16][k 4(l/16)+j] and B[k 4(l/16)+j][col l%16] (j = 0..3) and holds
// D[row 4(l/16)+j][col l%16].  A wavefront evaluates 16 cells per tile:
// 6 + 2 MFMAs per 16-column block of the 96->32 and 32->32 layers and 2 for
// the 32->9 layer.  Each lane holds its B operands (9.2 KB BF16 per wave in
// all) in registers for the whole grid-stride loop.
#include "seis3d_abc.hpp"

#include <algorithm>
#include <atomic>

namespace seis3d::abc {

namespace {

typedef short __attribute__((ext_vector_type(4))) bf16x4;
typedef float __attribute__((ext_vector_type(4))) floatx4;

constexpr int kWaves = 4;
constexpr int kBlock = 64 * kWaves;

// Round-to-nearest-even FP32 -> BF16 (finite inputs).
__device__ __forceinline__ unsigned short to_bf16(float f) {
    const unsigned int u = __float_as_uint(f);
    return static_cast<unsigned short>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float from_bf16(unsigned short h) { return __uint_as_float(static_cast<unsigned int>(h) << 16); }

// Face-frame component ch -> global field for normal axis `axis` (normal component
// first): each of the triples (vx, vy, vz), (sxx, syy, szz), (sxy, syz, sxz) rotates
// cyclically with the axis.
__host__ __device__ constexpr int face_field(int axis, int ch) { return 3 * (ch / 3) + (ch % 3 + axis) % 3; }
// Components odd in the normal coordinate (v_n, s_nt1, s_nt2) flip sign on low faces.
__host__ __device__ constexpr bool face_odd(int ch) { return ch == 0 || ch == 6 || ch == 8; }

static_assert(face_field(0, 5) == 5 && face_field(1, 0) == 1 && face_field(1, 5) == 3 && face_field(2, 6) == 8 &&
                  face_field(2, 8) == 7,
              "face frame rotation");

// Per-wave tile buffers; every phase of a tile touches only its own wave's slices.
struct Lds {
    // Rows are padded so that the 16 cells of a wave fall in different LDS banks;
    // BF16 rows keep the 8-byte alignment of the MFMA operand loads.
    unsigned short xs[kWaves][16][kIn + 8];
    unsigned short h1[kWaves][16][kHidden + 8];
    unsigned short h2[kWaves][16][kHidden + 8];
    float centre[kWaves][16][kOut];   // raw face-frame value of the cell itself (tap 3)
    float scale[kWaves][16];
    float imp[kWaves][16];            // impedance at the cell
};

// Orders this wave's LDS writes before its later LDS reads by other lanes.  The
// per-tile phases only touch the wave's own slices of the LDS, so waves never wait
// for each other inside the tile loops.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// MFMA B operands and biases of one lane (column m = lane % 16, k-quad q = lane / 16),
// read once from the weight image and held in registers for a kernel's lifetime.
struct RegWeights {
    bf16x4 w1[kIn / 16][2], w2[kHidden / 16][2], w3[kHidden / 16];
    float b1[2], b2[2], b3;
};

// Image layout: BF16 W1 [kHidden][kIn], W2 [kHidden][kHidden], W3 [kOutPad][kHidden],
// then FP32 b1 [kHidden], b2 [kHidden], b3 [kOutPad].  Every operand offset is a multiple
// of four BF16 values, so the 8-byte loads are aligned.
__device__ void load_weights(const void* image, RegWeights& W) {
    const unsigned short* w = static_cast<const unsigned short*>(image);
    const unsigned short* w2 = w + kHidden * kIn;
    const unsigned short* w3 = w2 + kHidden * kHidden;
    const float* b = reinterpret_cast<const float*>(w + kBf16Count);
    const int lane = int(threadIdx.x) & 63, m = lane & 15, q = lane >> 4;
#pragma unroll
    for (int kb = 0; kb < kIn / 16; ++kb)
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
            W.w1[kb][nb] = *reinterpret_cast<const bf16x4*>(&w[(nb * 16 + m) * kIn + kb * 16 + 4 * q]);
#pragma unroll
    for (int kb = 0; kb < kHidden / 16; ++kb) {
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
            W.w2[kb][nb] = *reinterpret_cast<const bf16x4*>(&w2[(nb * 16 + m) * kHidden + kb * 16 + 4 * q]);
        W.w3[kb] = *reinterpret_cast<const bf16x4*>(&w3[m * kHidden + kb * 16 + 4 * q]);
    }
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
        W.b1[nb] = b[nb * 16 + m];
        W.b2[nb] = b[kHidden + nb * 16 + m];
    }
    W.b3 = b[2 * kHidden + m];
}

// MLP of one 16-cell tile held in L.xs[wave]; returns lane's layer-3 outputs
// (rows 4q+j, column lane%16).  Optionally stores BF16-rounded activations.
__device__ floatx4 mlp_tile(Lds& L, const RegWeights& W, int wave, float* H1, float* H2, long long row0, int rows) {
    const int lane = int(threadIdx.x) & 63, m = lane & 15, q = lane >> 4;
    floatx4 acc[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
    for (int kb = 0; kb < kIn / 16; ++kb) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(&L.xs[wave][m][kb * 16 + 4 * q]);
#pragma unroll
        for (int nb = 0; nb < 2; ++nb) acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, W.w1[kb][nb], acc[nb], 0, 0, 0);
    }
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const unsigned short h = to_bf16(fmaxf(acc[nb][j] + W.b1[nb], 0.0f));
            L.h1[wave][4 * q + j][nb * 16 + m] = h;
            if (H1 && 4 * q + j < rows) H1[(row0 + 4 * q + j) * kHidden + nb * 16 + m] = from_bf16(h);
        }
    wave_sync();
    floatx4 acc2[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
    for (int kb = 0; kb < kHidden / 16; ++kb) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(&L.h1[wave][m][kb * 16 + 4 * q]);
#pragma unroll
        for (int nb = 0; nb < 2; ++nb)
            acc2[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, W.w2[kb][nb], acc2[nb], 0, 0, 0);
    }
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const unsigned short h = to_bf16(fmaxf(acc2[nb][j] + W.b2[nb], 0.0f));
            L.h2[wave][4 * q + j][nb * 16 + m] = h;
            if (H2 && 4 * q + j < rows) H2[(row0 + 4 * q + j) * kHidden + nb * 16 + m] = from_bf16(h);
        }
    wave_sync();
    floatx4 out = {0, 0, 0, 0};
#pragma unroll
    for (int kb = 0; kb < kHidden / 16; ++kb) {
        const bf16x4 a = *reinterpret_cast<const bf16x4*>(&L.h2[wave][m][kb * 16 + 4 * q]);
        out = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, W.w3[kb], out, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) out[j] += W.b3;
    return out;
}

__device__ __forceinline__ long long stride_of(const DevGrid& g, int axis) {
    return axis == 0 ? 1LL : axis == 1 ? static_cast<long long>(g.px) : g.plane;
}
__device__ __forceinline__ int coord_of(const DevGrid& g, long long i, int axis) {
    if (axis == 0) return int(i % g.px);
    if (axis == 1) return int((i / g.px) % g.gy);
    return int(i / g.plane);
}
__device__ __forceinline__ int extent_of(const DevGrid& g, int axis) { return axis == 0 ? g.gx : axis == 1 ? g.gy : g.gz; }

// Field c of a frame.  Selecting by a switch keeps the kernel-argument pointer
// array in scalar registers; indexing it with a per-lane value would copy it to
// scratch memory and add a scratch load to every feature load.
__device__ __forceinline__ const float* field_ptr(const Frame& F, int c) {
    switch (c) {
    case 0: return F.f[0];
    case 1: return F.f[1];
    case 2: return F.f[2];
    case 3: return F.f[3];
    case 4: return F.f[4];
    case 5: return F.f[5];
    case 6: return F.f[6];
    case 7: return F.f[7];
    default: return F.f[8];
    }
}

__device__ __forceinline__ float impedance(const Frame& F, long long i) {
    return sqrtf((F.lambda[i] + 2.0f * F.mu[i]) / F.inv_rho[i]);
}

// Features per lane: the 96 features (12 channels x 8 taps) of a cell are split over 4 lanes.
constexpr int kPerLane = kIn / 4;

// Raw (un-normalised) features part + 4k (k = 0..23) of frame cell i on a face normal to
// AXIS; returns the impedance at the cell.  Feature f is tap f % 8 of channel f / 8 (tap 3
// = the cell, taps run outwards along the normal).  With part < 4, feature part + 4k is
// channel k / 2, tap 4 (k & 1) + part, so once the loop is unrolled every channel's field,
// sign and scaling are compile-time constants, a lane's loads use two cell offsets, and
// all of them issue before the first result is used.
template <int AXIS>
__device__ __forceinline__ float gather_axis(const Frame& F, int i, bool hi, int part, float (&feat)[kPerLane]) {
    const long long st = stride_of(F.g, AXIS) * (hi ? 1 : -1);
    const int c0 = coord_of(F.g, i, AXIS), n = extent_of(F.g, AXIS);
    const float z = impedance(F, i);
    long long p[2];
    bool inside[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int o = 4 * h + part - 3;
        const int c = c0 + (hi ? o : -o);
        p[h] = i + o * st;
        inside[h] = c >= 4 && c < n - 4;
    }
#pragma unroll
    for (int k = 0; k < kPerLane; ++k) {
        const int ch = k / 2, h = k & 1;
        float x;
        if (ch < 9) {
            x = F.f[face_field(AXIS, ch)][p[h]];
            if (!hi && face_odd(ch)) x = -x;
            if (ch >= 3) x = x / z;
        } else if (ch == 9) {
            x = inside[h] ? 1.0f : 0.0f;
        } else if (ch == 10) {
            x = sqrtf((F.lambda[p[h]] + 2.0f * F.mu[p[h]]) * F.inv_rho[p[h]]) * (1.0f / 5000.0f);
        } else {
            x = inside[h] ? (hi ? 1.0f : -1.0f) : 0.0f;
        }
        feat[k] = x;
    }
    return z;
}

// Frame cells are listed in grid order, so the 16 cells of a tile nearly always share
// a normal axis and a wave takes one branch.  A constant axis keeps every field
// pointer a kernel-argument load (a lane-varying index into the pointer array would
// copy the whole argument block to scratch memory).
__device__ __forceinline__ float gather_features(const Frame& F, int i, unsigned char code, int part,
                                                 float (&feat)[kPerLane]) {
    const bool hi = (code & 1) != 0;
    switch (code >> 1) {
    case 0: return gather_axis<0>(F, i, hi, part, feat);
    case 1: return gather_axis<1>(F, i, hi, part, feat);
    default: return gather_axis<2>(F, i, hi, part, feat);
    }
}

// Inference: corrections of all frame cells into delta [count][9] (global field order).
// Lane (cell r = lane / 4, part = lane % 4) gathers and normalises its 24 features in
// registers; LDS holds only the BF16 tile, the activations and three values per cell.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void correct_kernel(Frame F, FrameCells C,
                                                                                                 const void* image,
                                                                                                 float* delta) {
    extern __shared__ unsigned char smem[];
    Lds& L = *reinterpret_cast<Lds*>(smem);
    RegWeights W;
    load_weights(image, W);
    const int wave = int(threadIdx.x) >> 6, lane = int(threadIdx.x) & 63;
    const int cellr = lane >> 2, part = lane & 3;
    const int tiles = (C.count + 15) / 16;
    const int waves_total = int(gridDim.x) * kWaves;
    const int iters = (tiles + waves_total - 1) / waves_total;
    for (int it = 0; it < iters; ++it) {
        const int tile = it * waves_total + int(blockIdx.x) * kWaves + wave;
        const int first = tile * 16;
        const int rows = tile < tiles ? min(16, C.count - first) : 0;
        float feat[kPerLane];
        float z = 0.0f;
        if (cellr < rows) {
            z = gather_features(F, C.cell[first + cellr], C.code[first + cellr], part, feat);
        } else {
#pragma unroll
            for (int k = 0; k < kPerLane; ++k) feat[k] = 0.0f;
        }
        // Scale s = max |field feature| of the cell (its first 72 features, k < 18 on each
        // of its 4 lanes).  A window quieter than kMinScale is silent: zero features and
        // zero scale, so no correction (1 / s would overflow for denormal s and turn zero
        // taps into NaN).
        float mx = 0.0f;
#pragma unroll
        for (int k = 0; k < 18; ++k) mx = fmaxf(mx, fabsf(feat[k]));
        mx = fmaxf(mx, __shfl_xor(mx, 1));
        mx = fmaxf(mx, __shfl_xor(mx, 2));
        const bool live = mx >= kMinScale;
        const float inv = live ? 1.0f / mx : 0.0f;
#pragma unroll
        for (int k = 0; k < kPerLane; ++k)
            L.xs[wave][cellr][part + 4 * k] = to_bf16(k < 18 ? feat[k] * inv : (live ? feat[k] : 0.0f));
        if (part == 3) {
#pragma unroll
            for (int ch = 0; ch < kOut; ++ch) L.centre[wave][cellr][ch] = feat[2 * ch];   // tap 3 of channel ch
        }
        if (part == 0) {
            L.scale[wave][cellr] = live ? mx : 0.0f;
            L.imp[wave][cellr] = z;
        }
        wave_sync();
        const floatx4 out = mlp_tile(L, W, wave, nullptr, nullptr, 0, 0);
        const int m = lane & 15, q = lane >> 4;
        if (m < kOut) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int r = 4 * q + j;
                if (r >= rows) continue;
                const unsigned char code = C.code[first + r];
                const int axis = code >> 1;
                const bool hi = (code & 1) != 0;
                // Dissipative correction: the network may only move a value towards zero
                // (never grow it or change its sign), so the frame can remove energy but
                // not add it.  Unrestricted one-step corrections feed back over long runs
                // and grow without bound.
                const float sc = L.scale[wave][r];
                float d = out[j];
                if (sc > 0.0f) {
                    const float x = L.centre[wave][r][m] / sc;
                    const float mag = fabsf(x);
                    const float kept = fminf(fmaxf(copysignf(1.0f, x) * (x + d), 0.0f), mag);
                    d = copysignf(kept, x) - x;
                } else {
                    d = 0.0f;
                }
                float v = d * sc;
                if (m >= 3) v *= L.imp[wave][r];
                if (!hi && face_odd(m)) v = -v;
                delta[static_cast<long long>(first + r) * kOut + face_field(axis, m)] = v;
            }
        }
        wave_sync();
    }
}

struct FieldPtrs {
    float* f[9];
};
__global__ void apply_kernel_v(FieldPtrs P, FrameCells C, const float* delta) {
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= C.count) return;
    const int i = C.cell[k];
#pragma unroll
    for (int c = 0; c < 9; ++c) P.f[c][i] += delta[static_cast<long long>(k) * kOut + c];
}

// Features, scales and face-frame scaled values of the listed cells (one thread per cell).
__global__ __launch_bounds__(128) void features_kernel(Frame F, FrameCells C, float* X, float* S, float* B) {
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= C.count) return;
    float row[kIn];
    const int i = C.cell[k];
    const unsigned char code = C.code[k];
#pragma unroll
    for (int part = 0; part < 4; ++part) {
        float feat[kPerLane];
        gather_features(F, i, code, part, feat);
#pragma unroll
        for (int j = 0; j < kPerLane; ++j) row[part + 4 * j] = feat[j];
    }
    float mx = 0.0f;
    for (int f = 0; f < 9 * kTaps; ++f) mx = fmaxf(mx, fabsf(row[f]));
    const bool live = mx >= kMinScale;
    const float inv = live ? 1.0f / mx : 0.0f;
    bool finite = true;
    for (int f = 0; f < kIn; ++f) {
        const float v = f < 9 * kTaps ? row[f] * inv : (live ? row[f] : 0.0f);
        finite = finite && isfinite(v);
        X[static_cast<long long>(k) * kIn + f] = v;
    }
    S[k] = live && finite ? mx : 0.0f;   // a zero scale keeps the row out of training samples
    for (int ch = 0; ch < 9; ++ch) B[static_cast<long long>(k) * kOut + ch] = row[ch * kTaps + 3];   // tap 3 = the cell
}

// Face-frame scaled values at the cells of grid `tg` mapped into the grid of R by (ox, oy, oz).
__global__ void face_values_kernel(Frame R, DevGrid tg, FrameCells C, int ox, int oy, int oz, float* out) {
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= C.count) return;
    const int i = C.cell[k];
    const unsigned char code = C.code[k];
    const int axis = code >> 1;
    const bool hi = (code & 1) != 0;
    const int x = coord_of(tg, i, 0), y = coord_of(tg, i, 1), z = coord_of(tg, i, 2);
    const long long p = (static_cast<long long>(z + oz) * R.g.gy + (y + oy)) * R.g.px + (x + ox);
    const float zimp = impedance(R, p);
    for (int ch = 0; ch < 9; ++ch) {
        float v = field_ptr(R, face_field(axis, ch))[p];
        if (!hi && face_odd(ch)) v = -v;
        if (ch >= 3) v = v / zimp;
        out[static_cast<long long>(k) * kOut + ch] = v;
    }
}

// Rows idx[k] of a [*][width] array into dst rows first + k.
__global__ void gather_rows_kernel(const float* src, const int* idx, int n, int width, float* dst, long long first) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long long>(n) * width) return;
    const long long k = t / width;
    const int c = int(t % width);
    dst[(first + k) * width + c] = src[static_cast<long long>(idx[k]) * width + c];
}

// Normalised targets of each row and their largest magnitude (infinity when a
// target is not finite), so the caller can drop rows the scale does not bound.
__global__ void targets_kernel(const float* V, const float* B, const float* S, int n, float* T, float* M) {
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= n) return;
    const float s = S[k];
    float mx = 0.0f;
    for (int c = 0; c < kOut; ++c) {
        const long long j = static_cast<long long>(k) * kOut + c;
        const float t = s > 0.0f ? (V[j] - B[j]) / s : 0.0f;
        T[j] = t;
        mx = isfinite(t) ? fmaxf(mx, fabsf(t)) : INFINITY;
    }
    M[k] = mx;
}

struct BoxArgs {
    DevGrid dst, src;
    float* df[9];
    const float* sf[9];
    int ox, oy, oz;
};
__global__ void copy_box_kernel(BoxArgs A) {
    const int x = int(blockIdx.x * blockDim.x + threadIdx.x), y = int(blockIdx.y), z = int(blockIdx.z);
    if (x >= A.dst.gx) return;
    const long long d = (static_cast<long long>(z) * A.dst.gy + y) * A.dst.px + x;
    const bool ghost = x < 4 || x >= A.dst.gx - 4 || y < 4 || y >= A.dst.gy - 4 || z < 4 || z >= A.dst.gz - 4;
    const long long s = (static_cast<long long>(z + A.oz) * A.src.gy + (y + A.oy)) * A.src.px + (x + A.ox);
#pragma unroll
    for (int c = 0; c < 9; ++c) A.df[c][d] = ghost ? 0.0f : A.sf[c][s];
}

// ---- training --------------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void forward_kernel(const void* image, const float* X, int n, float* H1, float* H2,
                                                         float* Y) {
    extern __shared__ unsigned char smem[];
    Lds& L = *reinterpret_cast<Lds*>(smem);
    RegWeights W;
    load_weights(image, W);
    const int wave = int(threadIdx.x) >> 6, lane = int(threadIdx.x) & 63;
    const int tiles = (n + 15) / 16;
    const int waves_total = int(gridDim.x) * kWaves;
    const int iters = (tiles + waves_total - 1) / waves_total;
    for (int it = 0; it < iters; ++it) {
        const int tile = it * waves_total + int(blockIdx.x) * kWaves + wave;
        const long long first = static_cast<long long>(tile) * 16;
        const int rows = tile < tiles ? int(min(16LL, n - first)) : 0;
        const int cellr = lane >> 2, part = lane & 3;
        for (int f = part; f < kIn; f += 4)
            L.xs[wave][cellr][f] = to_bf16(cellr < rows ? X[(first + cellr) * kIn + f] : 0.0f);
        wave_sync();
        const floatx4 out = mlp_tile(L, W, wave, H1, H2, first, rows);
        const int m = lane & 15, q = lane >> 4;
        if (m < kOut)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if (4 * q + j < rows) Y[(first + 4 * q + j) * kOut + m] = out[j];
        wave_sync();
    }
}

// One thread per sample: dY, dH2, dH1 through the BF16-rounded weights; squared error.
__global__ void backward_samples_kernel(const void* image, const float* H1, const float* H2, const float* Y,
                                        const float* T, int n, float denom, float* dY, float* dH2, float* dH1,
                                        float* err) {
    __shared__ float w2[kHidden * kHidden], w3[kOutPad * kHidden];
    const unsigned short* w = static_cast<const unsigned short*>(image);
    for (int k = threadIdx.x; k < kHidden * kHidden; k += blockDim.x) w2[k] = from_bf16(w[kHidden * kIn + k]);
    for (int k = threadIdx.x; k < kOutPad * kHidden; k += blockDim.x)
        w3[k] = from_bf16(w[kHidden * kIn + kHidden * kHidden + k]);
    __syncthreads();
    const long long s = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float dy[kOut], e = 0.0f;
    for (int o = 0; o < kOut; ++o) {
        const float r = Y[s * kOut + o] - T[s * kOut + o];
        e += r * r;
        dy[o] = 2.0f * r / denom;
        dY[s * kOut + o] = dy[o];
    }
    err[s] = e;
    float dh2[kHidden];
    for (int k = 0; k < kHidden; ++k) {
        float acc = 0.0f;
        for (int o = 0; o < kOut; ++o) acc += dy[o] * w3[o * kHidden + k];   // W3^T [o][k]
        dh2[k] = H2[s * kHidden + k] > 0.0f ? acc : 0.0f;
        dH2[s * kHidden + k] = dh2[k];
    }
    for (int i = 0; i < kHidden; ++i) {
        float acc = 0.0f;
        for (int k = 0; k < kHidden; ++k) acc += dh2[k] * w2[k * kHidden + i];   // W2^T [k][i]
        dH1[s * kHidden + i] = H1[s * kHidden + i] > 0.0f ? acc : 0.0f;
    }
}

// Partial parameter gradients over a contiguous chunk of samples; parameter p per thread loop.
__global__ void param_grads_kernel(const float* X, const float* H1, const float* H2, const float* dY, const float* dH2,
                                   const float* dH1, int n, int chunks, float* partial) {
    const int chunk = int(blockIdx.x);
    const long long per = (static_cast<long long>(n) + chunks - 1) / chunks;
    const long long s0 = chunk * per, s1 = min(static_cast<long long>(n), s0 + per);
    for (int p = threadIdx.x; p < kParams; p += blockDim.x) {
        float acc = 0.0f;
        if (p < kB1) {
            const int f = p / kHidden, i = p % kHidden;
            for (long long s = s0; s < s1; ++s) acc += from_bf16(to_bf16(X[s * kIn + f])) * dH1[s * kHidden + i];
        } else if (p < kW2) {
            const int i = p - kB1;
            for (long long s = s0; s < s1; ++s) acc += dH1[s * kHidden + i];
        } else if (p < kB2) {
            const int i = (p - kW2) / kHidden, k = (p - kW2) % kHidden;
            for (long long s = s0; s < s1; ++s) acc += H1[s * kHidden + i] * dH2[s * kHidden + k];
        } else if (p < kW3) {
            const int k = p - kB2;
            for (long long s = s0; s < s1; ++s) acc += dH2[s * kHidden + k];
        } else if (p < kB3) {
            const int k = (p - kW3) / kOut, o = (p - kW3) % kOut;
            for (long long s = s0; s < s1; ++s) acc += H2[s * kHidden + k] * dY[s * kOut + o];
        } else {
            const int o = p - kB3;
            for (long long s = s0; s < s1; ++s) acc += dY[s * kOut + o];
        }
        partial[static_cast<long long>(chunk) * kParams + p] = acc;
    }
}

__global__ void reduce_partials_kernel(const float* partial, int chunks, float* grad) {
    const int p = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= kParams) return;
    float acc = 0.0f;
    for (int c = 0; c < chunks; ++c) acc += partial[static_cast<long long>(c) * kParams + p];
    grad[p] = acc;
}

__global__ void pack_kernel(const float* p, unsigned short* image) {
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    float* bias = reinterpret_cast<float*>(image + kBf16Count);
    if (k < kHidden * kIn) {   // W1^T [j][i] = W1[i][j]
        const int j = k / kIn, i = k % kIn;
        image[k] = to_bf16(p[kW1 + i * kHidden + j]);
    } else if (k < kHidden * kIn + kHidden * kHidden) {
        const int r = k - kHidden * kIn, j = r / kHidden, i = r % kHidden;
        image[k] = to_bf16(p[kW2 + i * kHidden + j]);
    } else if (k < kBf16Count) {
        const int r = k - kHidden * kIn - kHidden * kHidden, o = r / kHidden, i = r % kHidden;
        image[k] = o < kOut ? to_bf16(p[kW3 + i * kOut + o]) : static_cast<unsigned short>(0);
    } else if (k < kBf16Count + kHidden) {
        bias[k - kBf16Count] = p[kB1 + (k - kBf16Count)];
    } else if (k < kBf16Count + 2 * kHidden) {
        bias[k - kBf16Count] = p[kB2 + (k - kBf16Count - kHidden)];
    } else if (k < kBf16Count + 2 * kHidden + kOutPad) {
        const int o = k - kBf16Count - 2 * kHidden;
        bias[k - kBf16Count] = o < kOut ? p[kB3 + o] : 0.0f;
    }
}

__global__ void adam_kernel(float* p, float* m, float* v, const float* g, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= kParams) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

__global__ void sum_partial_kernel(const float* x, int n, float* partial) {
    __shared__ float s[256];
    float acc = 0.0f;
    for (long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
         k += static_cast<long long>(gridDim.x) * blockDim.x)
        acc += x[k];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ void sum_final_kernel(const float* partial, int n, float* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        float acc = 0.0f;
        for (int k = 0; k < n; ++k) acc += partial[k];
        *out = acc;
    }
}

unsigned blocks_for(long long n, int per) { return unsigned(n > 0 ? (n + per - 1) / per : 1); }

constexpr int kSumBlocks = 256;

}  // namespace

// Blocks of correct_kernel resident on the whole device at once, cached per device: the
// grid-stride loop then runs every block for the same number of tiles, with no partial
// last round of blocks.
hipError_t resident_correct_blocks(int* blocks) {
    constexpr int kMaxDevices = 64;
    static std::atomic<int> cache[kMaxDevices];
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    const int cached = dev >= 0 && dev < kMaxDevices ? cache[dev].load(std::memory_order_relaxed) : 0;
    if (cached > 0) {
        *blocks = cached;
        return hipSuccess;
    }
    int per_cu = 0, cus = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, correct_kernel, kBlock, sizeof(Lds));
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return e;
    *blocks = std::max(1, per_cu) * std::max(1, cus);
    if (dev >= 0 && dev < kMaxDevices) cache[dev].store(*blocks, std::memory_order_relaxed);
    return hipSuccess;
}

hipError_t launch_correct(const Frame& F, const FrameCells& cells, const void* image, float* delta, hipStream_t s) {
    if (cells.count <= 0) return hipSuccess;
    int resident = 0;
    const hipError_t e = resident_correct_blocks(&resident);
    if (e != hipSuccess) return e;
    const int tiles = (cells.count + 15) / 16;
    const unsigned blocks = unsigned(std::min(resident, (tiles + kWaves - 1) / kWaves));
    hipLaunchKernelGGL(correct_kernel, dim3(blocks), dim3(kBlock), sizeof(Lds), s, F, cells, image, delta);
    return hipGetLastError();
}

hipError_t launch_apply(float* const f[9], const FrameCells& cells, const float* delta, hipStream_t s) {
    if (cells.count <= 0) return hipSuccess;
    FieldPtrs P;
    for (int c = 0; c < 9; ++c) P.f[c] = f[c];
    hipLaunchKernelGGL(apply_kernel_v, dim3(blocks_for(cells.count, 256)), dim3(256), 0, s, P, cells, delta);
    return hipGetLastError();
}

hipError_t launch_features(const Frame& F, const FrameCells& cells, float* X, float* S, float* B, hipStream_t s) {
    if (cells.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(features_kernel, dim3(blocks_for(cells.count, 128)), dim3(128), 0, s, F, cells, X, S, B);
    return hipGetLastError();
}

hipError_t launch_face_values(const Frame& R, const DevGrid& tgrid, const FrameCells& cells, int ox, int oy, int oz,
                              float* out, hipStream_t s) {
    if (cells.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(face_values_kernel, dim3(blocks_for(cells.count, 256)), dim3(256), 0, s, R, tgrid, cells, ox, oy,
                       oz, out);
    return hipGetLastError();
}

hipError_t launch_gather_rows(const float* src, const int* idx, int n, int width, float* dst, long long first,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_rows_kernel, dim3(blocks_for(static_cast<long long>(n) * width, 256)), dim3(256), 0, s,
                       src, idx, n, width, dst, first);
    return hipGetLastError();
}

hipError_t launch_targets(const float* V, const float* B, const float* S, int n, float* T, float* M, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(targets_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s, V, B, S, n, T, M);
    return hipGetLastError();
}

hipError_t launch_copy_box(const DevGrid& dst, float* const df[9], const DevGrid& src, const float* const sf[9],
                           int ox, int oy, int oz, hipStream_t s) {
    BoxArgs A{};
    A.dst = dst;
    A.src = src;
    for (int c = 0; c < 9; ++c) { A.df[c] = df[c]; A.sf[c] = sf[c]; }
    A.ox = ox; A.oy = oy; A.oz = oz;
    hipLaunchKernelGGL(copy_box_kernel, dim3(blocks_for(dst.gx, 64), unsigned(dst.gy), unsigned(dst.gz)), dim3(64), 0, s,
                       A);
    return hipGetLastError();
}

hipError_t launch_mlp_forward(const void* image, const float* X, int n, float* H1, float* H2, float* Y, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int tiles = (n + 15) / 16;
    const unsigned blocks = unsigned(std::min(2048, (tiles + kWaves - 1) / kWaves));
    hipLaunchKernelGGL(forward_kernel, dim3(blocks), dim3(kBlock), sizeof(Lds), s, image, X, n, H1, H2, Y);
    return hipGetLastError();
}

hipError_t launch_mlp_backward_samples(const void* image, const float* H1, const float* H2, const float* Y,
                                       const float* T, int n, float denom, float* dY, float* dH2, float* dH1,
                                       float* err, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(backward_samples_kernel, dim3(blocks_for(n, 128)), dim3(128), 0, s, image, H1, H2, Y, T, n,
                       denom, dY, dH2, dH1, err);
    return hipGetLastError();
}

hipError_t launch_mlp_param_grads(const float* X, const float* H1, const float* H2, const float* dY, const float* dH2,
                                  const float* dH1, int n, float* partial, int chunks, float* grad, hipStream_t s) {
    hipLaunchKernelGGL(param_grads_kernel, dim3(unsigned(chunks)), dim3(256), 0, s, X, H1, H2, dY, dH2, dH1, n, chunks,
                       partial);
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(blocks_for(kParams, 256)), dim3(256), 0, s, partial, chunks, grad);
    return hipGetLastError();
}

hipError_t launch_adam(float* p, float* m, float* v, const float* grad, float lr, float beta1, float beta2, float eps,
                       int step, void* image, hipStream_t s) {
    const float c1 = 1.0f - std::pow(beta1, float(step)), c2 = 1.0f - std::pow(beta2, float(step));
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(kParams, 256)), dim3(256), 0, s, p, m, v, grad, lr, beta1, beta2,
                       eps, c1, c2);
    return launch_pack(p, image, s);
}

hipError_t launch_pack(const float* p, void* image, hipStream_t s) {
    hipLaunchKernelGGL(pack_kernel, dim3(blocks_for(kBf16Count + 2 * kHidden + kOutPad, 256)), dim3(256), 0, s, p,
                       static_cast<unsigned short*>(image));
    return hipGetLastError();
}

hipError_t launch_sum(const float* x, int n, float* partial, float* out, hipStream_t s) {
    hipLaunchKernelGGL(sum_partial_kernel, dim3(kSumBlocks), dim3(256), 0, s, x, n, partial);
    hipLaunchKernelGGL(sum_final_kernel, dim3(1), dim3(64), 0, s, partial, kSumBlocks, out);
    return hipGetLastError();
}

}  // namespace seis3d::abc
