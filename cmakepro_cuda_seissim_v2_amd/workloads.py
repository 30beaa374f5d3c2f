"""Synthetic workloads of the benchmark configurations (seeded, no external data).

* ``layered_model``: 5 horizontal layers with seeded random interface depths,
  vp 1800..4500 m/s, vs = vp/1.73, Gardner density 310 * vp**0.25.
* ``gradient_random_model``: smooth vp 2000 -> 4000 m/s depth gradient with a seeded Gaussian
  random field (correlation length 20 cells, +-10 %).
* ``config2_context`` / ``config2_setup``: the 300 x 50 x 300 free-surface + 20-cell CPML run
  with a 15 Hz vertical force at (150, 25, 75) and the 150-receiver default spread.

Config 2 asks for dt = 1 ms at dx = 10 m with vp_max = 4500 m/s: Courant 0.45 exceeds the
O(8) 3D limit 0.44884 and validateContext3D (cuda/Cuda_Seis3D.cu:479-480) rejects it, so the
runs use the reference's automatic dt (847 us at CFL 0.85).
"""
from __future__ import annotations

import numpy as np

from . import seis3d as S

BASE_SEED = 20261203
LAYER_VP = (1800.0, 2475.0, 3150.0, 3825.0, 4500.0)


def gardner_density(vp: np.ndarray) -> np.ndarray:
    return (310.0 * np.power(vp, 0.25)).astype(np.float32)


def layered_model(nx: int, ny: int, nz: int, seed: int = BASE_SEED):
    """Five flat layers; interface depths drawn from a seeded generator."""
    rng = np.random.default_rng(seed)
    cuts = np.sort(rng.choice(np.arange(2, nz - 2), size=len(LAYER_VP) - 1, replace=False))
    layer = np.searchsorted(cuts, np.arange(nz), side="right")
    vp_z = np.asarray(LAYER_VP, np.float32)[layer]
    vp = np.broadcast_to(vp_z[:, None, None], (nz, ny, nx)).astype(np.float32)
    vs = (vp / np.float32(1.73)).astype(np.float32)
    return vp, vs, gardner_density(vp)


def gaussian_random_field(shape, corr: float, seed: int) -> np.ndarray:
    """FFT-filtered white noise with a Gaussian kernel of length `corr` cells, scaled to |g| <= 1."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape)
    k = [np.fft.fftfreq(n) for n in shape]
    kk = sum(np.meshgrid(*[ki ** 2 for ki in k], indexing="ij"))
    filt = np.exp(-0.5 * (2 * np.pi) ** 2 * kk * corr ** 2)
    g = np.real(np.fft.ifftn(np.fft.fftn(noise) * filt))
    return (g / np.max(np.abs(g))).astype(np.float32)


def gradient_random_model(nx: int, ny: int, nz: int, seed: int = BASE_SEED, corr: float = 20.0,
                          amplitude: float = 0.1):
    z = np.arange(nz, dtype=np.float32) / max(1, nz - 1)
    bg = (2000.0 + 2000.0 * z)[:, None, None]
    g = gaussian_random_field((nz, ny, nx), corr, seed)
    vp = (bg * (1.0 + amplitude * g)).astype(np.float32)
    vs = (vp / np.float32(1.73)).astype(np.float32)
    return vp, vs, gardner_density(vp)


def imported_context(nx: int, ny: int, nz: int, max_vp: float, pml: int = 20, free_surface: bool = True,
                     spacing: float = 10.0, frequency: float = 15.0, **kw) -> S.Ctx:
    c = S.default_context(nx=nx, ny=ny, nz=nz, pml=pml, free_surface=int(free_surface), spacing=spacing,
                          frequency=frequency, model_kind=S.IMPORTED, max_vp=float(max_vp), **kw)
    return S.validateContext3D(c)


def config2_context(nx: int = 300, ny: int = 50, nz: int = 300) -> S.Ctx:
    return imported_context(nx, ny, nz, max_vp=max(LAYER_VP))


def config2_source(ctx: S.Ctx) -> S.Source:
    """15 Hz vertical force at the reference's default source position (nx/2, ny/2, nz/4)."""
    return S.default_source(x=ctx.nx // 2, y=ctx.ny // 2, z=max(2, ctx.nz // 4), kind=S.FORCE,
                            amplitude=1.0e8, direction=(0.0, 0.0, 1.0))


def config2_receivers(ctx: S.Ctx):
    """The default 50 x 3 surface spread of fitSurvey3D/configureReceivers3D."""
    q = S.fitSurvey3D(ctx, S.default_survey())
    return S.configureReceivers3D(ctx, q)


def batch_shots(count: int = 16, seed: int = BASE_SEED, sizes=None, corr: float = 20.0, receivers: int = 100,
                pml: int = 20, frequency: float = 15.0):
    """Uneven multi-shot workload: per shot (context, (vp, vs, rho), source, receivers).

    nx, nz drawn from 128..256 in steps of 16 and ny = 128 unless `sizes` gives (nx, ny, nz) per shot;
    smooth vp 2000 -> 4000 m/s depth gradient times (1 + 0.1 g) with a seeded Gaussian random field
    g of correlation length `corr` cells; a seeded 15 Hz vertical force; a `receivers`-long surface
    line along x through the model's centre.
    """
    rng = np.random.default_rng(seed)
    shots = []
    for k in range(count):
        if sizes is not None:
            nx, ny, nz = sizes[k]
        else:
            nx = int(rng.choice(np.arange(128, 257, 16)))
            nz = int(rng.choice(np.arange(128, 257, 16)))
            ny = 128
        vp, vs, rho = gradient_random_model(nx, ny, nz, seed=seed + 1 + k, corr=corr)
        ctx = imported_context(nx, ny, nz, max_vp=float(vp.max()), pml=pml, frequency=frequency)
        sx = int(rng.integers(nx // 4, 3 * nx // 4))
        sz = int(rng.integers(4, max(5, nz // 4)))
        src = S.default_source(x=sx, y=ny // 2, z=sz, kind=S.FORCE, amplitude=1.0e8, direction=(0.0, 0.0, 1.0))
        nrec = min(receivers, nx)
        rec = [(int(round(i * (nx - 1) / max(1, nrec - 1))), ny // 2, 0) for i in range(nrec)]
        shots.append((ctx, (vp, vs, rho), src, rec))
    return shots


def fwi_workload(count: int = 8, nx: int = 200, ny: int = 100, nz: int = 200, seed: int = BASE_SEED,
                 radius: float = 15.0, gain: float = 0.08, receivers: int = 200, pml: int = 20,
                 frequency: float = 15.0, free_surface: bool = True):
    """FWI gradient workload: (context, start model, true model, [(source, receivers)] per shot).

    Background: smooth depth gradient vp 2000 -> 4000 m/s, vs = vp/1.73, Gardner density (the
    starting model). True model: the background with vp * (1 + gain * exp(-(r/radius)^2)) around
    a seeded centre in the middle third of the model. A free top with CPML on the other five
    faces (or CPML on all six). Shots: `count` 15 Hz explosions at depth 5 spread evenly along x
    at y = ny/2; each records `receivers` receivers on a line along x through y = ny/2, on the
    surface (depth 2 with CPML on top).
    """
    rng = np.random.default_rng(seed)
    z = np.arange(nz, dtype=np.float32) / max(1, nz - 1)
    vp = np.broadcast_to((2000.0 + 2000.0 * z)[:, None, None], (nz, ny, nx)).astype(np.float32)
    vs = (vp / np.float32(1.73)).astype(np.float32)
    start = (vp, vs, gardner_density(vp))
    cx, cy, cz = (float(rng.uniform(n / 3, 2 * n / 3)) for n in (nx, ny, nz))
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r2 = ((xx - cx) ** 2 + (yy - cy) ** 2 + (zz - cz) ** 2) / radius ** 2
    vp_true = (vp * (1.0 + gain * np.exp(-r2))).astype(np.float32)
    true = (vp_true, vs, gardner_density(vp))
    ctx = imported_context(nx, ny, nz, max_vp=float(vp_true.max()), pml=pml, free_surface=free_surface,
                           frequency=frequency)
    nrec = min(receivers, nx)
    depth = 0 if free_surface else 2
    rec = [(int(round(i * (nx - 1) / max(1, nrec - 1))), ny // 2, depth) for i in range(nrec)]
    shots = []
    for k in range(count):
        sx = int(round(nx * (k + 0.5) / count))
        shots.append((S.default_source(x=min(max(sx, 1), nx - 2), y=ny // 2, z=5, kind=S.EXPLOSION,
                                       amplitude=1.0e8), rec))
    return ctx, start, true, shots


def fwi_check_direction(ctx: S.Ctx, model, gradient, param: int, top: int = 20, ramp: int = 10) -> np.ndarray:
    """Central-difference direction for a full-size FWI gradient check along parameter `param`
    (0 vp, 1 vs, 2 rho): the gradient itself, scaled so that the largest relative change of that
    parameter is 1, zero where vp reaches the context's vp_max (uploads refuse a faster vp, and
    vp_max fixes dt), and ramped from zero at depth `top` to one at `top + ramp`.

    The ramp keeps the direction away from the explosions (depth 5) and the surface receivers:
    there the misfit is so strongly nonlinear in the model that central differences at any step
    FP32 misfits can resolve are dominated by their third-order term.  Below it they converge
    quadratically.
    """
    base = np.asarray(model[param], np.float64)
    vp = np.asarray(model[0], np.float64)
    taper = np.clip((ctx.max_vp - vp) / (0.05 * ctx.max_vp), 0.0, 1.0)
    depth = np.arange(base.shape[0], dtype=np.float64)[:, None, None]
    g = np.asarray(gradient, np.float64) * taper * np.clip((depth - top) / ramp, 0.0, 1.0)
    return g / float(np.max(np.abs(g))) * base
