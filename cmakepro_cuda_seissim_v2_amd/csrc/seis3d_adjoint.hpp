// seis3d_adjoint.hpp - host-visible launchers of the discrete-adjoint (FWI
// gradient) kernels of the 3D O(8) velocity-stress step.
#pragma once

#include <hip/hip_runtime.h>

#include "seis3d_device.hpp"

namespace seis3d {

// Operands of one adjoint step n, run backwards: the forward state it reads
// (stresses before step n, velocities and CPML memories after it), the model,
// and the adjoint state it updates.
struct AdjStep {
    DevGrid g;
    const float* fv[3];      // forward vx vy vz after step n
    const float* fs[6];      // forward sxx syy szz sxy syz sxz before step n
    const float* psi[3];     // forward CPML memories after step n
    const float* inv_rho;
    const float* lambda;
    const float* mu;
    const float* inv_mu;
    const float* coef[3];
    float* a[9];             // adjoint wavefields
    float* apsi[3];          // adjoint CPML memories (forward psi layout)
    float* t[9];             // filtered adjoint derivatives [3 v + axis]; zero outside the updated cells
    float* g_lam;            // sum_n dJ/dlambda at each cell (= dJ/dM at fixed mu, M = lambda + 2 mu)
    float* g_dev;            // sum_n dJ/dmu - 2 dJ/dlambda through the normal stresses (dJ/dmu at fixed M)
    float* g_shear[3];       // sum_n dJ/dmu_h of the xy, yz, xz harmonic shear moduli
    float* g_buoy[3];        // sum_n dJ/dh of the x, y, z face buoyancies
    float* surf[3];          // free surface: per-column scratch planes of the velocity-ghost transpose
    float dt;
    float invh;
    float spacing;
    int reflecting;          // reflecting walls (then the grid has no CPML)
    DevSource src;           // force source of step n (active only for a force source)
};

// Up to 16 same-kind array operations in one launch.
struct ArrayBatch {
    const float* src[16];
    float* dst[16];
    long long n[16];
    int count;
};

// Gradient finalisation: folds the padding onto the physical cells and
// applies the chain rule from (1/rho, lambda, mu) to (vp, vs, rho).
struct FoldParams {
    DevGrid g;
    int nx, ny, nz, ox, oy, oz;
    const float* inv_rho;
    const float* lambda;
    const float* mu;
    const float* g_ib;
    const float* g_lam;
    const float* g_dev;
    float* g_vp;             // [nz][ny][nx]
    float* g_vs;
    float* g_rho;
};

// Adjoint half-steps.  With a free surface or reflecting walls they include the
// transposes of the ghost fills: the stress half-step opens with the stress
// ghosts and closes with the velocity ghosts (surf holds three zeroed planes),
// the velocity half-step closes with the stress ghosts.
hipError_t launch_adjoint_stress(const AdjStep& P, hipStream_t s);
hipError_t launch_adjoint_velocity(const AdjStep& P, hipStream_t s);
// Adds R^T r of one receiver sample r ([receiver][vx, vy, vz]) to the adjoint velocities.
hipError_t launch_adjoint_inject(const AdjStep& P, const int* rec, int count, const float* residual, hipStream_t s);
// residual = gather - observed, elementwise.
hipError_t launch_residual(const float* gather, const float* observed, float* residual, long long n, hipStream_t s);
// Per-cell dJ/d(1/rho) into g_ib and the harmonic shear terms folded into out_mu.
hipError_t launch_adjoint_cells(const AdjStep& P, float* g_ib, float* out_dev, hipStream_t s);
hipError_t launch_adjoint_fold(const FoldParams& F, hipStream_t s);
hipError_t launch_copy_arrays(const ArrayBatch& b, hipStream_t s);
hipError_t launch_zero_arrays(const ArrayBatch& b, hipStream_t s);
// dst[k] += src[k]
hipError_t launch_add_arrays(const ArrayBatch& b, hipStream_t s);

}  // namespace seis3d
