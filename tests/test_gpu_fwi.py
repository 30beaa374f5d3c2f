"""FWI gradient on the GPU (seis3d_fwi_gradient): discrete adjoint with checkpoints and recompute.

No reference counterpart. Every case runs with all-CPML boundaries, with a free top (whose ghost
fills and surface law have their own transposes) and with reflecting walls. Pinned by
* the FP64 and FP32 oracle adjoints on the same seeded cases (the FP64 one is itself pinned by
  central differences in test_oracle_adjoint.py);
* central differences of the GPU's own forward misfit along seeded directions;
* invariances: the checkpoint interval (1, 7, all steps) changes nothing, bit for bit; the
  out-of-place recompute of the LDS-DMA kernels equals the copy route of the in-place kernel
  families bit for bit; a batched request equals the sum of single-shot requests; the
  session's own wavefield is untouched.
"""
import numpy as np
import pytest

import fwi_cases as F
from conftest import hip_device_count
from cmakepro_cuda_seissim_v2_amd import seis3d as S
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

NX, NY, NZ, STEPS = 18, 16, 20, 70


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def observed_data(ctx, true, src, rec, steps, stride):
    with O.Oracle3D(ctx, 64) as o:
        o.upload_model(*true)
        o.set_receivers(rec)
        return o.run(src, 0, steps, stride=stride, max_samples=steps // stride).astype(np.float32)


def gpu_gradient(ctx, model, shots, steps, stride, interval=0):
    with S.GPUSimData3D(ctx) as g:
        g.upload_model(*model)
        return g.fwi_gradient(shots, steps, stride, interval)


def gpu_misfit(ctx, model, shots, steps, stride):
    """0.5 |gather - observed|^2 from plain forward runs with the gather API."""
    total = 0.0
    for src, rec, obs in shots:
        with S.GPUSimData3D(ctx) as g:
            g.upload_model(*model)
            g.set_receivers(rec)
            g.gather_begin(steps // stride, stride)
            g.run(src, 0, steps)
            got = g.gather_read()
        total += 0.5 * float(np.sum((got.astype(np.float64) - obs.astype(np.float64)) ** 2))
    return total


@pytest.fixture(scope="module", params=[(S.EXPLOSION, 1, "cpml"), (S.FORCE, 2, "cpml"), (S.EXPLOSION, 1, "free"),
                                        (S.FORCE, 2, "free"), (S.EXPLOSION, 1, "walls"), (S.FORCE, 2, "walls")],
                ids=["explosion", "force-stride2", "explosion-free", "force-stride2-free", "explosion-walls",
                     "force-stride2-walls"])
def small(request):
    kind, stride, boundary = request.param
    start, true = F.models(NX, NY, NZ)
    ctx = F.context(NX, NY, NZ, true, free_surface=boundary == "free", reflecting=boundary == "walls")
    src = F.source(kind, NX, NY, NZ)
    rec = F.receivers(NX, NY, NZ, surface=boundary == "free")
    obs = observed_data(ctx, true, src, rec, STEPS, stride)
    return dict(ctx=ctx, start=start, shots=[(src, rec, obs)], stride=stride)


def oracle_gradient(case, precision):
    (src, rec, obs), = case["shots"]
    with O.Oracle3D(case["ctx"], precision) as o:
        o.upload_model(*case["start"])
        o.set_receivers(rec)
        return o.fwi_gradient(src, STEPS, case["stride"], obs.astype(np.float64))


def test_gradient_matches_oracle(gpu, small):
    misfit, gvp, gvs, grho = gpu_gradient(small["ctx"], small["start"], small["shots"], STEPS, small["stride"], 10)
    for precision, tol in ((64, 1e-4), (32, 1e-4)):
        m_ref, *g_ref = oracle_gradient(small, precision)
        assert misfit == pytest.approx(m_ref, rel=1e-4)
        errs = [rel_l2(a, b) for a, b in zip((gvp, gvs, grho), g_ref)]
        print(f"FP{precision} oracle: misfit {misfit:.6e} vs {m_ref:.6e}, rel-L2 vp/vs/rho {errs}")
        assert max(errs) < tol, (precision, errs)


def test_misfit_matches_forward_runs(gpu, small):
    misfit, *_ = gpu_gradient(small["ctx"], small["start"], small["shots"], STEPS, small["stride"])
    assert misfit == pytest.approx(gpu_misfit(small["ctx"], small["start"], small["shots"], STEPS, small["stride"]),
                                   rel=1e-6)


def test_checkpoint_interval_is_bitwise_neutral(gpu, small):
    ref = gpu_gradient(small["ctx"], small["start"], small["shots"], STEPS, small["stride"], STEPS)
    for interval in (1, 7, 0):
        got = gpu_gradient(small["ctx"], small["start"], small["shots"], STEPS, small["stride"], interval)
        assert got[0] == ref[0]
        for a, b in zip(got[1:], ref[1:]):
            assert np.array_equal(a, b), interval


@pytest.mark.parametrize("family,tile_y,z_chunk,interval", [(0, 4, 16, 7), (0, 8, 32, 70), (0, 4, 16, 9)])
def test_out_of_place_recompute_equals_copy_path(gpu, small, family, tile_y, z_chunk, interval):
    """The LDS-DMA kernels step recomputed segments out of place, straight into the stored
    states; the in-place direct family steps a live state and copies it after every step.
    The two families are bitwise-identical, so both routes give the same gradient bit for bit.
    Interval 9 leaves a short last segment (70 = 7 * 9 + 7), whose unused slots the
    out-of-place route must clear itself."""
    results = []
    for fam, ty, zc in ((4, 4, 16), (family, tile_y, z_chunk)):
        with S.GPUSimData3D(small["ctx"]) as g:
            g.upload_model(*small["start"])
            g.set_tiles(fam, ty, zc)
            results.append(g.fwi_gradient(small["shots"], STEPS, small["stride"], interval))
    assert results[0][0] == results[1][0]
    for a, b in zip(results[0][1:], results[1][1:]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("boundary", ["cpml", "free", "walls"])
def test_batched_request_equals_single_shots(gpu, boundary):
    """A request with several shots (on their own streams, in lockstep) must equal the sum of the
    shots requested one at a time, with CPML all round, a free top and reflecting walls."""
    start, true = F.models(NX, NY, NZ)
    ctx = F.context(NX, NY, NZ, true, free_surface=boundary == "free", reflecting=boundary == "walls")
    rec = F.receivers(NX, NY, NZ, surface=boundary == "free")
    shots = []
    for k, kind in enumerate((S.EXPLOSION, S.FORCE, S.EXPLOSION)):
        src = F.source(kind, NX, NY, NZ, shift=3 * k - 3)
        shots.append((src, rec, observed_data(ctx, true, src, rec, STEPS, 1)))
    batched = gpu_gradient(ctx, start, shots, STEPS, 1, 9)
    singles = [gpu_gradient(ctx, start, [s], STEPS, 1, 9) for s in shots]
    assert batched[0] == pytest.approx(sum(s[0] for s in singles), rel=1e-12)
    for k in range(1, 4):
        total = np.sum([s[k].astype(np.float64) for s in singles], axis=0)
        assert rel_l2(batched[k], total) < 1e-6


def test_session_wavefield_untouched(gpu, small):
    (src, rec, obs), = small["shots"]
    with S.GPUSimData3D(small["ctx"]) as g:
        g.upload_model(*small["start"])
        g.set_receivers(rec)
        g.run(src, 0, 25)
        before = [g.download_field(f) for f in (0, 5, 8)]
        g.fwi_gradient(small["shots"], STEPS, small["stride"], 10)
        after = [g.download_field(f) for f in (0, 5, 8)]
        assert all(np.array_equal(a, b) for a, b in zip(before, after))
        g.run(src, 25, 5)   # the session still steps


# ---- central differences of the GPU forward misfit -----------------------------------------------

FDX, FDY, FDZ, FD_STEPS = 40, 24, 40, 160


@pytest.fixture(scope="module", params=["cpml", "free", "walls"])
def medium(request):
    if hip_device_count() < 1:
        pytest.skip("no HIP device")
    start, true = F.models(FDX, FDY, FDZ, radius=5.0)
    ctx = F.context(FDX, FDY, FDZ, true, frequency=30.0, free_surface=request.param == "free",
                    reflecting=request.param == "walls")
    rec = F.receivers(FDX, FDY, FDZ, surface=request.param == "free")
    shots = []
    for k, kind in enumerate((S.EXPLOSION, S.FORCE)):
        src = F.source(kind, FDX, FDY, FDZ, shift=6 * k - 3)
        shots.append((src, rec, observed_data(ctx, true, src, rec, FD_STEPS, 2)))
    grad = gpu_gradient(ctx, start, shots, FD_STEPS, 2, 20)
    return dict(ctx=ctx, start=start, shots=shots, grad=grad)


# (parameter mask, seed): vp, vs, rho alone, then vp and all three together.
FD_DIRECTIONS = [((1, 0, 0), 31), ((0, 1, 0), 32), ((0, 0, 1), 33), ((1, 0, 0), 34), ((1, 1, 1), 35)]


@pytest.mark.parametrize("mask,seed", FD_DIRECTIONS)
def test_gradient_matches_gpu_finite_differences(gpu, medium, mask, seed):
    """Richardson-extrapolated central differences (steps 0.4 % and 0.8 %) within 1e-3."""
    base = [np.asarray(a, np.float32) for a in medium["start"]]
    d = [F.smooth_direction(base[0].shape, seed + 100 * k, base[k], 1.0) * mask[k] for k in range(3)]

    def central(eps):
        js, moved = [], []
        for sign in (1, -1):
            m = [(b.astype(np.float64) + sign * eps * dk).astype(np.float32) for b, dk in zip(base, d)]
            moved.append(m)
            js.append(gpu_misfit(medium["ctx"], tuple(m), medium["shots"], FD_STEPS, 2))
        fd = js[0] - js[1]
        pred = sum(float(np.sum(g.astype(np.float64) * (p.astype(np.float64) - q.astype(np.float64))))
                   for g, p, q in zip(medium["grad"][1:], moved[0], moved[1]))
        return fd, pred

    fd1, p1 = central(4e-3)
    fd2, p2 = central(8e-3)
    # Each difference is compared with the adjoint's prediction for its own float32 step;
    # the ratio removes the step size, Richardson removes the second-order truncation term.
    r1, r2 = fd1 / p1, fd2 / p2
    ratio = (4 * r1 - r2) / 3
    print(f"direction {mask} seed {seed}: fd/adjoint {r1:.6f} (0.4 %), {r2:.6f} (0.8 %), extrapolated {ratio:.6f}")
    assert abs(ratio - 1) < 1e-3, (r1, r2, ratio)
