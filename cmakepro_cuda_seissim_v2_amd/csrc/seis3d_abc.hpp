// seis3d_abc.hpp - learned absorbing boundary (no reference counterpart):
// an 8-cell frame whose cells take the plain (CPML-free) update; the outer 4
// layers, whose stencils reach the zero ghost cells, then take a correction
// from a pointwise MLP, 96 inputs -> 32 -> 32 -> 9, with BF16 weights on
// MFMA and FP32 accumulation.
//
// Inputs of a corrected cell: 8 taps along its outward normal (3 inward, the
// cell, 4 outward) x 12 channels: the 9 wavefields in the face frame
// (normal component first, signs mirrored on low faces, stresses divided by
// the cell's impedance rho vp), an inside-the-grid mask, vp / 5000 and the
// face side.  Field channels are divided by their largest magnitude s; the
// 9 outputs are face-frame corrections times s, so the correction is
// positively homogeneous in the wavefield and zero where the field is zero.
#pragma once

#include <hip/hip_runtime.h>

#include "seis3d_device.hpp"

namespace seis3d::abc {

constexpr int kIn = 96, kHidden = 32, kOut = 9, kOutPad = 16, kTaps = 8, kChannels = 12;
// Windows whose largest |field feature| is below this are silent (no normalisation, no correction).
constexpr float kMinScale = 1e-30f;
// FP32 packing of the parameters: W1 [96][32], b1 [32], W2 [32][32], b2 [32], W3 [32][9], b3 [9].
constexpr int kW1 = 0, kB1 = kW1 + kIn * kHidden, kW2 = kB1 + kHidden, kB2 = kW2 + kHidden * kHidden,
              kW3 = kB2 + kHidden, kB3 = kW3 + kHidden * kOut, kParams = kB3 + kOut;
// Device BF16 image: W1^T [32][96], W2^T [32][32], W3^T [16][32] (rows 9..15 zero), then FP32 biases 32 + 32 + 16.
constexpr int kBf16Count = kHidden * kIn + kHidden * kHidden + kOutPad * kHidden;
constexpr int kImageBytes = kBf16Count * 2 + (kHidden + kHidden + kOutPad) * 4;

// One frame cell: pitched index and (axis << 1 | high side).
struct FrameCells {
    const int* cell;
    const unsigned char* code;
    int count;
};

struct Frame {
    DevGrid g;
    const float* f[9];          // wavefields (read)
    const float* inv_rho;
    const float* lambda;
    const float* mu;
};

hipError_t launch_correct(const Frame& F, const FrameCells& cells, const void* image, float* delta, hipStream_t s);
hipError_t launch_apply(float* const f[9], const FrameCells& cells, const float* delta, hipStream_t s);
// Normalised inputs X [n][96], scales S [n] and face-frame scaled values B [n][9] of the listed cells.
hipError_t launch_features(const Frame& F, const FrameCells& cells, float* X, float* S, float* B, hipStream_t s);
// Face-frame scaled values [n][9] at the listed cells of grid `tgrid`, read from grid R at (x+ox, y+oy, z+oz).
hipError_t launch_face_values(const Frame& R, const DevGrid& tgrid, const FrameCells& cells, int ox, int oy, int oz,
                              float* out, hipStream_t s);
// dst rows first.. = src rows idx[0..n) of width floats.
hipError_t launch_gather_rows(const float* src, const int* idx, int n, int width, float* dst, long long first,
                              hipStream_t s);
// Normalised targets T = (V - B) / S (0 where S = 0) and M = max |T| per row
// (infinity for a row with a non-finite target).
hipError_t launch_targets(const float* V, const float* B, const float* S, int n, float* T, float* M, hipStream_t s);
// Copy of the non-ghost cells of one grid from another at a cell offset; ghost cells cleared.
hipError_t launch_copy_box(const DevGrid& dst, float* const df[9], const DevGrid& src, const float* const sf[9],
                           int ox, int oy, int oz, hipStream_t s);

// Training: forward over a batch (X rows first..first+n), storing BF16-rounded activations.
hipError_t launch_mlp_forward(const void* image, const float* X, int n, float* H1, float* H2, float* Y, hipStream_t s);
// Per-sample backward: dY = 2 (Y - T) / denom, dH2, dH1; squared errors per sample into err.
hipError_t launch_mlp_backward_samples(const void* image, const float* H1, const float* H2, const float* Y,
                                       const float* T, int n, float denom, float* dY, float* dH2, float* dH1,
                                       float* err, hipStream_t s);
// Parameter gradients (FP32 packing) in `chunks` partial sums, then reduced in order into grad.
hipError_t launch_mlp_param_grads(const float* X, const float* H1, const float* H2, const float* dY, const float* dH2,
                                  const float* dH1, int n, float* partial, int chunks, float* grad, hipStream_t s);
// Adam step on FP32 parameters, then the BF16 image.
hipError_t launch_adam(float* p, float* m, float* v, const float* grad, float lr, float beta1, float beta2, float eps,
                       int step, void* image, hipStream_t s);
hipError_t launch_pack(const float* p, void* image, hipStream_t s);
// Sum of n floats into one (in order of fixed partial blocks).
hipError_t launch_sum(const float* x, int n, float* partial, float* out, hipStream_t s);

}  // namespace seis3d::abc
