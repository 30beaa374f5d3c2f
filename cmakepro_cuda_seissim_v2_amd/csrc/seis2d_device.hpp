// seis2d_device.hpp - device operands of the 2D P-SV kernels.
//
// HBM layout: every field and material is one float array of nx * nz cells,
// row-major [z][x] with flat index k = z * nx + x (the reference's
// indexing, Cuda_Check.cu:229, so receiver and source indices carry over).
// A 400 x 400 model holds 14-20 such arrays (9-13 MB), which stay resident
// in the MI355X's L2/MALL between steps; the step is launch- and
// latency-bound, so a chunk of steps is captured once as a HIP graph.
#pragma once

#include <hip/hip_runtime.h>

namespace seis2d {

struct Dev2D {
    int nx, nz, n;
    int fs;               // free-surface row (= npml)
    int up;               // free surface on
    int type;             // 1 split PML, 2 hybrid, 3 free surface
    float dt, inv_dt;
    float c1, c2, c3, c4; // O(8) weights / h
    // 0 vx1 1 vx2 2 vz1 3 vz2 4 sxx1 5 sxx2 6 szz1 7 szz2 8 sxz1 9 sxz2 10 vx 11 vz 12 sxx 13 szz 14 sxz
    float* f[15];
    const float* rho;
    const float* mu;
    const float* lambda;
    const float* l2m;
    const float* dp;      // free-surface modulus per column (type 3)
    const float* dx;
    const float* dxh;
    const float* dz;
    const float* dzh;
    const float* mu_x;    // type-2 face averages (Cuda_Check.cu:924-947)
    const float* lam_x;
    const float* l2m_x;
    const float* mu_z;
    const float* rho_xz;
    const float* rho_o;
    int* counter;         // step index of graph-replayed runs (null: per-call steps)
};

// One step's single source (applyPointSource, Cuda_Check.cu:567-636).
struct Src2D {
    int active;
    int idx;              // flat index, free-surface row guard applied
    int kind;             // src_type
    float value;          // wavelet[it] * amp * 1e7 when `table` is null
    const float* table;   // graph runs: per-step values indexed by *counter; NaN = no injection
    float sin_a, cos_a;   // rotated force, evaluated on the host
};

struct MultiSrc {
    int idx;
    float value;          // amp * ricker * 1e7, evaluated on the host
};

struct Rec2D {
    const int* idx;
    int count;
    float* vx;            // per-step record ([count]) or traces ([count][len])
    float* vz;
    int len;              // trace length (0: per-step record)
    int first;            // step index of trace column 0
};

// The plane-sized work that follows a velocity half-step (applyPointSource or
// injectMultipleSourcesKernel, zeroBoundaryKernel, extractReceiverSamplesKernel),
// run by one workgroup.
struct Tail2D {
    Src2D single;            // the single source (active = 0: none)
    const MultiSrc* list;    // multi-source mode: list[count] (count = 0: off)
    int count;
    int kind;
    float sin_a, cos_a;
    Rec2D rec;               // rec.count = 0: no receiver samples this step
};

// Stress half-step (type 3 with the surface on: then the ghost-row mirror).
hipError_t launch_stress2d(const Dev2D& g, hipStream_t s);
// Velocity half-step, then the tail.
hipError_t launch_velocity2d(const Dev2D& g, const Tail2D& tail, hipStream_t s);
hipError_t launch_receivers2d(const Dev2D& g, const Rec2D& r, hipStream_t s);

}  // namespace seis2d
