"""Learned CPML profile on the GPU (seis3d_set_cpml_profile; no reference counterpart).

The reference's CPML tables come from one fixed profile (makeCoefficients3D,
cuda/Cuda_Seis3D.cu:45-64).  A learned profile changes only the tables, so:
  * the session's device tables equal the host computation and the oracle's restatement;
  * a learned-profile session steps bit for bit like the FP32 oracle with the same profile
    (the step kernels are the plain CPML ones) and within 1e-4 of the FP64 oracle;
  * None restores the reference tables bit for bit, and a CPML refresh keeps the profile;
  * the calls are refused while a batch holds the session or a learned frame is attached;
  * the shipped profile reflects less energy than the reference profile on models, sources and
    frequencies it was not fitted on (seeds outside its training set): the learned boundary
    beats its judge, 8-cell CPML, at the same per-step cost.
"""
import numpy as np
import pytest

from cmakepro_cuda_seissim_v2_amd import learned_abc as LA
from cmakepro_cuda_seissim_v2_amd import seis3d as S
from cmakepro_cuda_seissim_v2_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

PROFILES = [(-0.45, 3.6, 1.6, 0.9, 2.0), (0.4, 2.0, 6.0, 0.3, 0.5)]


def log(msg):
    print(msg, flush=True)


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-300))


def context(free_surface=True, **kw):
    base = dict(nx=40, ny=24, nz=32, max_vp=4500.0, pml=8, frequency=22.0, free_surface=free_surface)
    base.update(kw)
    return W.imported_context(base.pop("nx"), base.pop("ny"), base.pop("nz"), **base)


@pytest.mark.parametrize("free_surface", [True, False])
@pytest.mark.parametrize("profile", PROFILES)
def test_learned_profile_steps_like_the_oracle(gpu, free_surface, profile):
    c = context(free_surface)
    model = W.layered_model(c.nx, c.ny, c.nz, seed=5)
    src = S.default_source(x=8, y=12, z=10, kind=S.EXPLOSION, amplitude=1e8)
    rec = [(2 + 3 * i, 12, 0) for i in range(12)]
    steps, stride = 150, 2
    g = S.GPUSimData3D(c)
    g.upload_model(*model)
    g.set_cpml_profile(profile)
    p, learned = g.cpml_profile()
    assert learned and np.array_equal(p, np.asarray(profile))
    for axis in range(3):
        dev = g.read_cpml(axis)
        assert np.array_equal(dev, S.cpml_profile_coefficients(c, profile, axis)), axis
        assert np.array_equal(dev, O.profile_coefficients(c, profile, axis)), axis
    g.set_receivers(rec)
    g.gather_begin(steps // stride, stride)
    g.run(src, 0, steps)
    gather = g.gather_read().astype(np.float64)
    snap = [g.download_field(f) for f in range(9)]
    g.release()
    refs = {}
    for precision in (32, 64):
        with O.Oracle3D(c, precision) as o:
            o.upload_model(*model)
            o.set_profile(profile)
            o.set_receivers(rec)
            refs[precision] = (o.run(src, 0, steps, stride=stride, max_samples=steps // stride),
                               [o.field(f) for f in range(9)])
    assert np.any(gather)
    e32 = rel_l2(gather, refs[32][0])
    e64 = rel_l2(gather, refs[64][0])
    s32 = max(rel_l2(snap[f], refs[32][1][f]) for f in range(9))
    s64 = max(rel_l2(snap[f], refs[64][1][f]) for f in range(9))
    log(f"profile {profile} free {free_surface}: gather {e32:.2e} / {e64:.2e}, snapshots {s32:.2e} / {s64:.2e}")
    assert e32 == 0.0 and s32 == 0.0
    assert e64 < 1e-4 and s64 < 1e-4


def test_profile_reset_and_refresh(gpu):
    c = context(frequency=14.0)
    g = S.GPUSimData3D(c)
    reference = [g.read_cpml(axis) for axis in range(3)]
    g.set_cpml_profile(PROFILES[0])
    assert not np.array_equal(g.read_cpml(0), reference[0])
    new = c.copy()
    new.frequency = 27.0
    g.refresh_cpml(new)    # keeps the learned profile, at the new frequency
    for axis in range(3):
        assert np.array_equal(g.read_cpml(axis), S.cpml_profile_coefficients(new, PROFILES[0], axis))
    g.refresh_cpml(c)
    g.set_cpml_profile(None)
    p, learned = g.cpml_profile()
    assert not learned and tuple(p) == S.REFERENCE_PROFILE
    for axis in range(3):
        assert np.array_equal(g.read_cpml(axis).view(np.uint32), reference[axis].view(np.uint32))
    # The reference profile given explicitly gives the reference tables here too.
    g.set_cpml_profile(S.REFERENCE_PROFILE)
    for axis in range(3):
        assert np.array_equal(g.read_cpml(axis), reference[axis])
    with pytest.raises(S.InvalidArgument):
        g.set_cpml_profile((0.0, 0.5, 3.0, 1.0, 1.0))
    p, learned = g.cpml_profile()
    assert learned and tuple(p) == S.REFERENCE_PROFILE   # a rejected profile changes nothing
    g.release()


def test_profile_refused_while_batched_or_framed(gpu):
    c = context()
    g = S.GPUSimData3D(c)
    g.upload_model(*W.layered_model(c.nx, c.ny, c.nz, seed=3))
    batch = S.Batch([g], [W.config2_source(c)], 10)
    with pytest.raises(S.LogicError):
        g.set_cpml_profile(PROFILES[0])
    batch.close()
    g.abc_attach(None)
    with pytest.raises(S.LogicError):
        g.set_cpml_profile(PROFILES[0])
    g.abc_detach()
    g.set_cpml_profile(PROFILES[0])
    g.release()


@pytest.mark.timeout(600)
def test_shipped_profile_beats_reference_cpml_on_held_out_cases(gpu):
    profile = LA.load_profile()
    gains = []
    for seed in LA.PROFILE_HELD_OUT_SEEDS[:3]:
        assert seed not in LA.PROFILE_TRAIN_SEEDS
        case = LA.profile_case(seed, size=(48, 48, 48), steps=300, every=6)
        ref = LA.profile_reflection(None, case)
        got = LA.profile_reflection(profile, case)
        log(f"held-out seed {seed} ({case['frequency']:.1f} Hz): reference profile {ref:.3e}, learned {got:.3e}, "
            f"gain {ref / got:.2f}")
        gains.append(ref / got)
    assert min(gains) > 1.5


def test_fwi_gradient_with_learned_profile_matches_oracle(gpu):
    """The adjoint reads the session's CPML tables, so a learned profile carries through the FWI
    gradient: it matches the FP32 and FP64 oracle adjoints with the same profile, and differs from
    the gradient with the reference profile by far more than from the FP32 oracle."""
    import fwi_cases as F
    nx, ny, nz, steps = 18, 16, 20, 160
    learned = PROFILES[1]
    start, true = F.models(nx, ny, nz)
    ctx = F.context(nx, ny, nz, true, free_surface=True)
    src = F.source(S.EXPLOSION, nx, ny, nz)
    rec = F.receivers(nx, ny, nz, surface=True)
    with O.Oracle3D(ctx, 64) as o:
        o.upload_model(*true)
        o.set_profile(learned)
        o.set_receivers(rec)
        obs = o.run(src, 0, steps, stride=1, max_samples=steps).astype(np.float32)
    grads = {}
    for name, profile in (("learned", learned), ("reference", None)):
        with S.GPUSimData3D(ctx) as g:
            g.upload_model(*start)
            g.set_cpml_profile(profile)
            grads[name] = g.fwi_gradient([(src, rec, obs)], steps, 1, 10)
    misfit, *grad = grads["learned"]
    err32 = 0.0
    for precision in (32, 64):
        with O.Oracle3D(ctx, precision) as o:
            o.upload_model(*start)
            o.set_profile(learned)
            o.set_receivers(rec)
            m_ref, *g_ref = o.fwi_gradient(src, steps, 1, obs.astype(np.float64))
        errs = [rel_l2(a, b) for a, b in zip(grad, g_ref)]
        log(f"learned-profile FWI vs FP{precision} oracle: misfit {misfit:.6e} / {m_ref:.6e}, rel-L2 {errs}")
        assert misfit == pytest.approx(m_ref, rel=1e-4)
        assert max(errs) < 1e-4
        if precision == 32:
            err32 = max(errs)
    shift = min(rel_l2(a, b) for a, b in zip(grads["reference"][1:], grad))
    log(f"reference-profile gradient differs by {shift:.3e} (FP32 oracle: {err32:.3e})")
    assert shift > 30 * err32
