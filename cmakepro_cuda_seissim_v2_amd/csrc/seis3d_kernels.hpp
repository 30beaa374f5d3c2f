// seis3d_kernels.hpp - host-visible launchers of the 3D kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "seis3d_device.hpp"

namespace seis3d {

// Kernel family and tile shape of the two fused half-steps.
struct KernelConfig {
    int family = 4;    // 0 direct (one thread per cell), 4 LDS-DMA 2.5D stream
    int tile_y = 4;    // rows per workgroup of family 4: 4 or 8
    int z_chunk = 16;  // planes per workgroup of family 4
};

// Procedural model parameters of generateMaterialsKernel3D.
struct MaterialParams {
    int nx, ny, nz, ox, oy, oz;
    int kind;
    float vp, vs, rho, layer_vp, layer_vs, layer_rho;
    float interface_depth, layer_dip, layer_thickness, fault_throw, anomaly_radius;
};

// One packed slice: p[fixed_axis] = fixed, p[u_axis] = u + u_origin,
// p[v_axis] = v + v_origin in physical coordinates, then + (ox, oy, oz).
struct SliceParams {
    int fixed_axis, fixed, u_axis, v_axis;
    int u_origin, v_origin, width, height;
    int ox, oy, oz;
    int component;
};

// Packed views (makeViewLayout3D / extractViews3D, cuda/Cuda_Seis3D.cu:370-418,486-501).
struct ViewParams {
    long long slice_offset[3], volume_offset, total;
    int volume[3], domain[3], origin[3];
    int slices[3];             // physical x, y, z of the three planes
    int nx, ny, nz;            // physical extent (statistics cover physical cells unless include_boundary)
    int ox, oy, oz;            // padded index of physical (0, 0, 0)
    int component, material_component, include_boundary;
};
// stats: [0] max |wave|, [1] non-finite flag, [2] material min, [3] material max, [4] max |v|
// (float bits of non-negative values, so unsigned ordering is float ordering).
hipError_t launch_extract_views(const DevStep& P, const ViewParams& vp, long long count, float* out, unsigned* stats,
                                hipStream_t s);

hipError_t launch_velocity(const DevStep& P, const KernelConfig& k, hipStream_t s);
// LDS-DMA staged 2.5D family (seis3d_dma.hip), tile_y 4 or 8.
hipError_t launch_dma(bool velocity, const DevStep& P, int tile_y, int zchunk, hipStream_t s);
// Tile map of the LDS-DMA family for one grid.
TileMap dma_tile_map(const DevGrid& g, int tile_y, int zchunk);
// One LDS-DMA half-step (in place) of the L.nshots shots of L, one launch; all share tile_y.
hipError_t launch_dma_shots(bool velocity, const ShotLaunch& L, int tile_y, hipStream_t s);
// Work-queue half-step (64x4 tiles, in place) over the tiles of nshots shots: steps and maps are
// device arrays of nshots, first holds nshots + 1 item offsets, xcd_first the 9 item offsets of
// the eight XCDs' ranges (xcd_first[8] == first[nshots]), counters eight zeroed words; `blocks` is
// a whole number of rounds of eight (queue_grid: every workgroup resident).  Each shot's z-chunk
// length (its TileMap's) is at most 128 planes.
hipError_t launch_dma_queue(bool velocity, const DevStep* steps, const TileMap* maps, const int* first,
                            const int* xcd_first, unsigned* counters, int nshots, int blocks, hipStream_t s);
hipError_t queue_grid(int* blocks);
// Zero n words (graph-safe: a kernel, not a memset node).
hipError_t launch_zero_words(unsigned* w, long long n, hipStream_t s);
// Resident workgroups of the family-4 pair per CU (the smaller of velocity and
// stress) and the device's CU count, for sizing launches in whole rounds.
hipError_t dma_residency(int tile_y, int* blocks_per_cu, int* cus);
hipError_t launch_stress(const DevStep& P, const KernelConfig& k, hipStream_t s);
hipError_t launch_stress_ghost(const DevStep& P, hipStream_t s);
hipError_t launch_velocity_ghost(const DevStep& P, float spacing, hipStream_t s);
hipError_t launch_reflecting_walls(const DevStep& P, bool velocities, hipStream_t s);
// One receiver sample of one shot within a multi-shot launch.
struct ShotSample {
    int shot;              // index into the launch's step table
    int count;             // receivers
    const int* rec;        // padded coordinates, 3 ints each
    float* out;            // [receiver][vx, vy, vz]
};
// Velocity ghosts of every free-surface shot in `steps` (device array of nshots), one launch.
hipError_t launch_velocity_ghost_shots(const DevStep* steps, const float* spacing, int nshots, int max_gx, int max_gy,
                                       hipStream_t s);
// The receiver samples `samples` (device array of `entries`) of several shots, one launch.
hipError_t launch_receivers_shots(const DevStep* steps, const ShotSample* samples, int entries, int max_count,
                                  hipStream_t s);
hipError_t launch_receivers(const DevStep& P, const int* rec, int count, float* out, hipStream_t s);
hipError_t launch_generate_materials(const DevStep& P, const MaterialParams& mp, hipStream_t s);
hipError_t launch_extend_materials(const DevStep& P, const MaterialParams& mp, hipStream_t s);
hipError_t launch_reciprocal_mu(const DevStep& P, hipStream_t s);
hipError_t launch_extract_slice(const DevStep& P, const SliceParams& sp, float* out, hipStream_t s);

}  // namespace seis3d
