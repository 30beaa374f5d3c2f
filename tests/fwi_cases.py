"""Seeded FWI test cases shared by the oracle-adjoint and GPU-gradient tests.

A smooth gradient model with a weak random field is the starting model; the true model adds a
Gaussian anomaly (+8 % vp, -4 % rho) in the middle. All-CPML boundaries unless a case asks for a
free top or reflecting walls; receivers on two lines near the bottom, the source near the top.
"""
import numpy as np

from cmakepro_cuda_seissim_v2_amd import seis3d as S
from cmakepro_cuda_seissim_v2_amd import workloads as W


def models(nx, ny, nz, seed=2, radius=3.0, vp_gain=0.08, rho_gain=-0.04):
    vp, vs, rho = W.gradient_random_model(nx, ny, nz, seed=seed, corr=3.0, amplitude=0.05)
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    blob = np.exp(-((xx - nx / 2) ** 2 + (yy - ny / 2) ** 2 + (zz - nz / 2) ** 2) / (2 * radius ** 2))
    vp_true = (vp * (1 + vp_gain * blob)).astype(np.float32)
    rho_true = (rho * (1 + rho_gain * blob)).astype(np.float32)
    return (vp, vs, rho), (vp_true, vs, rho_true)


def context(nx, ny, nz, true_model, pml=8, frequency=40.0, free_surface=False, reflecting=False):
    return W.imported_context(nx, ny, nz, max_vp=float(true_model[0].max() * 1.05), pml=pml,
                              free_surface=free_surface and not reflecting, frequency=frequency,
                              reflecting=int(reflecting))


def source(kind, nx, ny, nz, shift=0):
    if kind == S.EXPLOSION:
        return S.default_source(x=nx // 2 + shift, y=ny // 2, z=3, kind=S.EXPLOSION, amplitude=1e8)
    return S.default_source(x=nx // 2 - 2 + shift, y=ny // 2, z=4, kind=S.FORCE, amplitude=1e8,
                            direction=(0.6, 0.0, 0.8))


def receivers(nx, ny, nz, surface=False):
    """Two lines near the bottom; with `surface`, two more on the free surface, whose vz samples
    reach half into the ghost plane above it."""
    rec = [(x, y, nz - 2) for y in (3, ny - 4) for x in range(2, nx - 2, 3)]
    if surface:
        rec += [(x, y, 0) for y in (4, ny - 5) for x in range(3, nx - 3, 4)]
    return rec


def smooth_direction(shape, seed, base, scale):
    """A seeded smooth perturbation of relative size `scale` of the array `base`."""
    return W.gaussian_random_field(shape, 3.0, seed).astype(np.float64) * np.asarray(base, np.float64) * scale
