"""Batched multi-shot runs and the tile autotuner on the GPU.

There is no reference counterpart (the reference runs shots one after another,
src/Seis3D_Sim.cpp:133-143); what pins the batch is that every shot's gather equals, bit for bit,
the same shot run on its own with the reference-style loop (runStep3D + sampleReceivers3D every
stride), and for the autotuner that its choice gives the same wavefield as the default kernels.
"""
import numpy as np
import pytest

from cmakepro_cuda_seissim_v2_amd import seis3d as S
from cmakepro_cuda_seissim_v2_amd import workloads as W

pytestmark = pytest.mark.gpu


def prepare(ctx, model, rec, tiles=None):
    g = S.GPUSimData3D(ctx)
    if tiles:
        g.set_tiles(*tiles)
    g.upload_model(*model)
    g.set_receivers(rec)
    return g


def plain_shot(ctx, model, src, rec, steps, stride):
    """The reference's per-shot loop: one runStep3D per step, sampleReceivers3D every stride."""
    g = prepare(ctx, model, rec)
    samples = []
    for n in range(steps):
        g.step(src, n)
        if (n + 1) % stride == 0:
            samples.append(g.sample_receivers())
    g.release()
    return np.stack(samples)


@pytest.mark.parametrize("mode", ["shot_streams", "shared_launches", "work_queue"])
def test_batched_shots_equal_single_shots(gpu, mode):
    shots = W.batch_shots(5, seed=11, sizes=[(40, 24, 36), (56, 20, 32), (33, 27, 41), (48, 32, 24), (36, 36, 36)],
                          corr=4.0, receivers=20, pml=8)
    steps, stride = 120, 2
    sessions = []
    for ctx, model, src, rec in shots:
        g = prepare(ctx, model, rec)
        g.gather_begin(steps // stride, stride)
        sessions.append(g)
    batch = S.Batch(sessions, [s[2] for s in shots], steps, mode=mode)
    batch.launch()
    batch.synchronize()
    for (ctx, model, src, rec), g in zip(shots, sessions):
        got = g.gather_read()
        ref = plain_shot(ctx, model, src, rec, steps, stride)
        assert got.shape == ref.shape == (steps // stride, len(rec), 3)
        assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))
    batch.close()
    for g in sessions:
        g.release()


@pytest.mark.parametrize("mode", ["shot_streams", "work_queue"])
def test_batch_replay_is_deterministic(gpu, mode):
    """Every replay of the graph restarts from the reset wavefields (and, in the work-queue mode,
    from freshly zeroed queues), with eager work and host synchronisation between replays."""
    shots = W.batch_shots(2, seed=3, sizes=[(40, 24, 40), (36, 28, 32)], corr=4.0, receivers=10, pml=8)
    sessions = []
    for ctx, model, src, rec in shots:
        g = prepare(ctx, model, rec)
        g.gather_begin(50, 1)
        sessions.append(g)
    batch = S.Batch(sessions, [s[2] for s in shots], 50, mode=mode)
    out = []
    for _ in range(3):
        for g in sessions:
            g.reset()
            g.synchronize()
        batch.launch()
        batch.synchronize()
        out.append([g.gather_read() for g in sessions])
    for later in out[1:]:
        assert all(np.array_equal(a, b) for a, b in zip(out[0], later))
    batch.close()
    for g in sessions:
        g.release()


def test_autotune_choice_is_valid_and_bitwise_neutral(gpu):
    ctx, model, src, rec = W.batch_shots(1, seed=5, sizes=[(64, 32, 48)], corr=4.0, receivers=10, pml=10)[0]
    g = prepare(ctx, model, rec)
    best = g.autotune(src, steps=3)
    assert best["family"] in (0, 4) and best["ms_per_step"] > 0
    assert g.get_tiles() == (best["family"], best["tile_y"], best["z_chunk"])
    g.run(src, 0, 40)
    h = prepare(ctx, model, rec, tiles=(0, 8, 32))
    h.run(src, 0, 40)
    for f in range(9):
        assert np.array_equal(g.download_field(f), h.download_field(f))
    g.release()
    h.release()


@pytest.mark.parametrize("mode", ["shared_launches", "work_queue"])
def test_mixed_boundaries_and_sources_equal_single_shots(gpu, mode):
    """Shots that differ in everything a multi-shot launch carries per shot: CPML thickness, a free
    top or CPML on all six faces, explosion or oblique force sources, receiver counts."""
    base = W.batch_shots(4, seed=23, sizes=[(40, 24, 36), (44, 28, 30), (36, 32, 40), (48, 20, 28)], corr=4.0,
                         receivers=12, pml=8)
    shots = []
    for k, (ctx, model, src, rec) in enumerate(base):
        vp = model[0]
        c = W.imported_context(ctx.nx, ctx.ny, ctx.nz, max_vp=float(vp.max()), pml=8 + 2 * k,
                               free_surface=k % 2 == 0)
        if k % 2:
            s = S.default_source(x=ctx.nx // 2, y=ctx.ny // 2, z=ctx.nz // 2, kind=S.EXPLOSION, amplitude=1.0e8)
            r = [(x, ctx.ny // 2, 3) for x in range(2, ctx.nx - 2, 4)]
        else:
            s = S.default_source(x=ctx.nx // 3, y=ctx.ny // 2, z=4, kind=S.FORCE, amplitude=1.0e8,
                                 direction=(0.6, 0.0, 0.8))
            r = rec
        shots.append((c, model, s, r))
    steps, stride = 90, 3
    sessions = []
    for ctx, model, src, rec in shots:
        g = prepare(ctx, model, rec)
        g.gather_begin(steps // stride, stride)
        sessions.append(g)
    batch = S.Batch(sessions, [s[2] for s in shots], steps, mode=mode)
    batch.launch()
    batch.synchronize()
    for (ctx, model, src, rec), g in zip(shots, sessions):
        got = g.gather_read()
        ref = plain_shot(ctx, model, src, rec, steps, stride)
        assert np.any(ref)
        assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))
    batch.close()
    for g in sessions:
        g.release()
