"""GPU parity of the HIP solver against the CPU oracle, through the C ABI.

Bar (north star): receiver indices and sample times bit-exact; gathers and wavefield
snapshots within rel-L2 1e-5 of the FP32 oracle (the restatement of the reference FP32 path)
and 1e-4 of the FP64 oracle.  Every assertion message states the measured error.
"""
import sys

import numpy as np
import pytest

from cmakepro_cuda_seissim_v2_amd import seis3d as S
from cmakepro_cuda_seissim_v2_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-4


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = {
    # name: (ctx builder, source, model builder or None, steps, stride)
    "free-top-layered-force": (
        lambda: W.imported_context(40, 24, 36, max_vp=4500.0, pml=8),
        dict(x=20, y=12, z=9, kind=S.FORCE, direction=(0.0, 0.0, 1.0)),
        lambda c: W.layered_model(c.nx, c.ny, c.nz, seed=3), 160, 2),
    "all-cpml-twolayer-explosion": (
        lambda: S.validateContext3D(S.default_context(nx=37, ny=19, nz=23, pml=9, free_surface=0,
                                                      model_kind=S.TWO_LAYER, frequency=25.0)),
        dict(x=18, y=9, z=11, kind=S.EXPLOSION), None, 120, 1),
    "free-top-gradient-oblique-force": (
        lambda: W.imported_context(30, 26, 28, max_vp=4400.0, pml=10, frequency=20.0),
        dict(x=11, y=14, z=6, kind=S.FORCE, direction=(0.48, -0.6, 0.64)),
        lambda c: W.gradient_random_model(c.nx, c.ny, c.nz, seed=4, corr=4.0), 140, 3),
    # The reference's limits at once: the smallest grid (16 per axis), a source at the smallest
    # valid position next to the free surface (x = y = 1, z = 2; validateContext3D / :604-607).
    "min-grid-edge-explosion": (
        lambda: W.imported_context(16, 16, 16, max_vp=4500.0, pml=8, frequency=30.0),
        dict(x=1, y=1, z=2, kind=S.EXPLOSION),
        lambda c: W.layered_model(c.nx, c.ny, c.nz, seed=6), 100, 1),
    # The thickest CPML (40 cells), the largest CFL (0.9) and a force at the far corner.
    "max-pml-cfl-corner-force": (
        lambda: S.validateContext3D(S.default_context(nx=20, ny=18, nz=24, pml=40, free_surface=0, cfl=0.9,
                                                      model_kind=S.DIPPING_LAYER, frequency=20.0)),
        dict(x=18, y=16, z=22, kind=S.FORCE, direction=(-0.6, 0.48, 0.64)), None, 100, 2),
    "reflecting-uniform-explosion": (
        lambda: S.validateContext3D(S.default_context(nx=24, ny=20, nz=22, pml=12, reflecting=1,
                                                      model_kind=S.UNIFORM)),
        dict(x=12, y=10, z=11, kind=S.EXPLOSION), None, 100, 2),
}

TILES = [(0, 8, 32), (4, 4, 5), (4, 4, 16), (4, 4, 37), (4, 8, 7), (4, 8, 16), (4, 8, 128), (4, 8, 0)]


def receivers_for(c):
    rec = [(x, y, 0) for y in (1, c.ny // 2, c.ny - 2) for x in range(1, c.nx - 1, 3)]
    rec += [(c.nx // 3, c.ny // 3, c.nz // 2), (c.nx - 1, c.ny - 1, c.nz - 1), (0, 0, 0)]
    return rec[:300]


def run_oracle(c, src, model, steps, stride, precision, rec):
    o = O.Oracle3D(c, precision)
    if model is not None:
        o.upload_model(*model)
    o.set_receivers(rec)
    g = o.run(src, 0, steps, stride=stride, max_samples=steps // stride)
    return o, g


def run_gpu(c, src, model, steps, stride, rec, tiles):
    g = S.GPUSimData3D(c)
    g.set_tiles(*tiles)
    if model is not None:
        g.upload_model(*model)
    g.set_receivers(rec)
    g.gather_begin(steps // stride, stride)
    g.run(src, 0, steps)
    gather = g.gather_read()
    return g, gather


@pytest.fixture(scope="module")
def oracle_runs():
    cache = {}

    def get(name, precision):
        key = (name, precision)
        if key not in cache:
            build, sk, mb, steps, stride = CASES[name]
            c = build()
            src = S.default_source(**sk)
            model = mb(c) if mb else None
            rec = receivers_for(c)
            o, gth = run_oracle(c, src, model, steps, stride, precision, rec)
            fields = [o.field(f) for f in range(9)]
            cache[key] = (gth, fields)
        return cache[key]
    return get


@pytest.mark.parametrize("tiles", TILES, ids=lambda t: f"fam{t[0]}-y{t[1]}-z{t[2]}")
@pytest.mark.parametrize("name", list(CASES))
def test_gather_and_snapshot_parity(gpu, oracle_runs, name, tiles):
    build, sk, mb, steps, stride = CASES[name]
    c = build()
    src = S.default_source(**sk)
    model = mb(c) if mb else None
    rec = receivers_for(c)
    log(f"case {name} tiles {tiles} dims {c.gx}x{c.gy}x{c.gz} steps {steps}")
    g, gather = run_gpu(c, src, model, steps, stride, rec, tiles)
    g32, f32 = oracle_runs(name, 32)
    g64, f64 = oracle_runs(name, 64)
    # Receiver indices and sample times: every sample slot filled, count exact.
    assert gather.shape == (steps // stride, len(rec), 3) == g32.shape
    e32 = rel_l2(gather, g32)
    e64 = rel_l2(gather, g64)
    exact = int(np.sum(gather.astype(np.float32) == g32.astype(np.float32)))
    log(f"  gather rel-L2 vs fp32 oracle {e32:.3e} ({exact}/{gather.size} bit-equal), vs fp64 {e64:.3e}")
    assert e32 <= TOL32, f"gather rel-L2 vs FP32 oracle {e32:.3e}"
    assert e64 <= TOL64, f"gather rel-L2 vs FP64 oracle {e64:.3e}"
    # Snapshots of the nine wavefields over the updated region (ghost planes included for v).
    core = (slice(4, c.gz - 4), slice(4, c.gy - 4), slice(4, c.gx - 4))
    for f in range(9):
        snap = g.download_field(f)
        s32 = rel_l2(snap[core], f32[f][core])
        s64 = rel_l2(snap[core], f64[f][core])
        assert s32 <= TOL32, f"field {f} rel-L2 vs FP32 oracle {s32:.3e}"
        assert s64 <= TOL64, f"field {f} rel-L2 vs FP64 oracle {s64:.3e}"
    g.release()


GENERATED = [S.UNIFORM, S.TWO_LAYER, S.SPHERE, S.THREE_LAYER, S.DIPPING_LAYER, S.LOW_VELOCITY_LAYER, S.FAULT,
             S.GRADIENT, S.SALT_DOME]


def generated_context(kind, free_surface=1, **kw):
    """A generated model (generateMaterialsKernel3D, cuda/Cuda_Seis3D.cu:114-147) with non-default
    interface, dip, thickness, throw and radius so every branch of its kind is exercised."""
    return S.validateContext3D(S.default_context(
        nx=40, ny=30, nz=36, pml=8, free_surface=free_surface, model_kind=kind, vp=2800.0, vs=1600.0,
        density=2100.0, layer_vp=3900.0, layer_vs=2200.0, layer_density=2450.0, interface_depth=0.45,
        layer_dip=0.3, layer_thickness=0.2, fault_throw=0.15, anomaly_radius=0.27, **kw))


@pytest.mark.parametrize("kind", GENERATED)
def test_generated_materials_bit_exact(gpu, kind):
    """1/rho, lambda and mu of every generated model kind, padding (clamp extension) included,
    equal the oracle's bit for bit: both sides use the same elementary operations without
    contraction."""
    c = generated_context(kind)
    g = S.GPUSimData3D(c)
    with O.Oracle3D(c, 32) as o:
        for field in (S.INV_RHO, S.LAMBDA, S.MU):
            got = g.download_field(field)
            want = o.field(field).astype(np.float32)
            diff = int(np.sum(got != want))
            assert diff == 0, f"model kind {kind} field {field}: {diff} cells differ"
    g.release()


@pytest.mark.parametrize("kind", [S.FAULT, S.SALT_DOME])
def test_generated_model_gather_parity(gpu, kind):
    """A force shot through the fault and salt-dome models: gather within the FP32 tolerance."""
    c = generated_context(kind, frequency=25.0)
    src = S.default_source(x=20, y=15, z=4, kind=S.FORCE, direction=(0.3, 0.0, 0.95))
    rec = receivers_for(c)
    steps, stride = 120, 2
    g, gather = run_gpu(c, src, None, steps, stride, rec, (4, 4, 16))
    o, g32 = run_oracle(c, src, None, steps, stride, 32, rec)
    e32 = rel_l2(gather, g32)
    log(f"model kind {kind}: gather rel-L2 vs fp32 oracle {e32:.3e}")
    assert gather.shape == g32.shape and e32 <= TOL32, f"gather rel-L2 vs FP32 oracle {e32:.3e}"
    g.release()


def test_stream_equals_direct_bitwise(gpu):
    """Both kernel families round identically (same operations, no contraction)."""
    c = W.imported_context(70, 30, 44, max_vp=4500.0, pml=10)
    model = W.layered_model(c.nx, c.ny, c.nz, seed=9)
    src = W.config2_source(c)
    out = []
    for tiles in [(0, 8, 32), (4, 4, 7), (4, 4, 16), (4, 8, 20), (4, 8, 128), (4, 8, 0)]:
        g = S.GPUSimData3D(c)
        g.set_tiles(*tiles)
        g.upload_model(*model)
        g.run(src, 0, 60)
        out.append([g.download_field(f) for f in range(9)])
        g.release()
    for f in range(9):
        for k in range(1, len(out)):
            assert np.array_equal(out[0][f], out[k][f]), (f, k)


@pytest.mark.parametrize("grid", [(300, 300, 400), (128, 16, 200)], ids=["no-round-fits", "thin-y"])
def test_balanced_shape_where_one_round_does_not_fit(gpu, grid):
    """The balanced one-round request (64x8 tiles, z_chunk 0) on a grid whose columns do not fit
    one round falls back to uniform chunks, and on a grid with few columns caps its chunks at 8
    planes; either way it steps the same bits as the default shape."""
    c = W.imported_context(*grid, max_vp=4500.0, pml=20)
    model = W.layered_model(c.nx, c.ny, c.nz, seed=4)
    src = W.config2_source(c)
    out = []
    for tiles in [(4, 4, 16), (4, 8, 0)]:
        g = S.GPUSimData3D(c)
        g.set_tiles(*tiles)
        g.upload_model(*model)
        g.run(src, 0, 24)
        out.append([g.download_field(f) for f in range(9)])
        g.release()
    assert np.any(out[0][2])
    for f in range(9):
        assert np.array_equal(out[0][f], out[1][f]), f


def test_balanced_shape_random_grids(gpu):
    """Seeded grids of assorted sizes, CPML widths and boundaries (all-CPML, free top, reflecting
    walls): the balanced one-round shape steps the same bits as the default shape."""
    rng = np.random.default_rng(17)
    for case in range(10):
        nx, ny, nz = (int(v) for v in rng.integers(16, 140, 3))
        pml = int(rng.integers(8, 25))
        mode = case % 3
        c = W.imported_context(nx, ny, nz, max_vp=4500.0, pml=pml, free_surface=mode == 1,
                               reflecting=int(mode == 2))
        model = W.layered_model(c.nx, c.ny, c.nz, seed=case)
        src = S.default_source(x=nx // 2, y=ny // 2, z=nz // 3, kind=S.EXPLOSION if case % 2 else S.FORCE,
                               amplitude=1e8, direction=(0.3, 0.4, 0.866))
        out = []
        for tiles in [(4, 4, 16), (4, 8, 0)]:
            g = S.GPUSimData3D(c)
            g.set_tiles(*tiles)
            g.upload_model(*model)
            g.run(src, 0, 40)
            out.append([g.download_field(f) for f in range(9)])
            g.release()
        print(f"case {case}: {nx}x{ny}x{nz} pml {pml} mode {mode}", flush=True)
        for f in range(9):
            assert np.array_equal(out[0][f], out[1][f]), (case, f)
        assert np.any(out[0][3])
