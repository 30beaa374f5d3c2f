"""Learned 8-cell absorbing boundary on the GPU (no reference counterpart).

* The MLP kernels (BF16 MFMA forward, FP32 backward) against the CPU oracle: outputs within the
  BF16 rounding of the oracle's forward, the parameter gradient against the oracle's backward
  on the same rounded activations to 1e-4.
* The frame: attaching without parameters and detaching restores plain 8-cell CPML bit for bit;
  a network whose output layer is zero changes nothing bit for bit.
* Training on the GPU lowers the loss.
* With the trained weights shipped in the package, the learned frame reflects less energy than
  the same frame without its network (truncation); plain 8- and 20-cell CPML are reported beside.
* Over long runs the dissipative correction keeps the learned frame bounded.
"""
import numpy as np
import pytest

from cmakepro_cuda_seissim_v2_amd import learned_abc as LA
from cmakepro_cuda_seissim_v2_amd import seis3d as S
from cmakepro_cuda_seissim_v2_amd import workloads as W
from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def random_case(n=1001, seed=0):
    rng = np.random.default_rng(seed)
    X = (rng.standard_normal((n, 96)) * 0.5).astype(np.float32)
    T = (rng.standard_normal((n, 9)) * 0.1).astype(np.float32)
    p = LA.init_params(seed)
    p[LA.N_PARAMS - 297:] = (rng.standard_normal(297) * 0.2).astype(np.float32)
    p[96 * 32:96 * 32 + 32] = (rng.standard_normal(32) * 0.1).astype(np.float32)
    return p, X, T


def test_mlp_forward_matches_oracle(gpu):
    p, X, T = random_case()
    Y, H1, H2, _, _ = LA.mlp(p, X, T)
    y, h1, h2, _ = M.forward(p, X)
    assert rel(Y, y) < 2e-3, rel(Y, y)
    # Hidden activations are BF16 values; FP32 (MFMA) vs FP64 sums flip a rounding only rarely.
    assert np.mean(H1 == h1) > 0.99 and np.mean(H2 == h2) > 0.98
    assert np.array_equal(H1, M.bf16(H1)) and np.array_equal(H2, M.bf16(H2))


def test_mlp_gradient_matches_oracle(gpu):
    p, X, T = random_case(seed=1)
    Y, H1, H2, g, loss = LA.mlp(p, X, T)
    ref_loss, ref_g, ref_y = M.loss_and_grad(p, X, T, h1=H1, h2=H2)
    assert loss == pytest.approx(ref_loss, rel=1e-4)
    assert rel(Y, ref_y) < 1e-5
    assert rel(g, ref_g) < 1e-4, rel(g, ref_g)
    parts = LA.unpack(g)
    ref_parts = LA.unpack(ref_g.astype(np.float32))
    for k in parts:
        assert rel(parts[k], ref_parts[k]) < 1e-3, k


def small_model(n=40, seed=3):
    return LA.layered_random_model(n, n, n, seed)


def run_fields(g, src, steps):
    for k in range(steps):
        g.step(src, k)
    return [g.download_field(f) for f in (0, 2, 5)]


def test_frame_attach_detach_is_exact(gpu):
    model = small_model()
    ctx = W.imported_context(40, 40, 40, max_vp=float(model[0].max()), pml=8, free_surface=True, frequency=25.0)
    src = S.default_source(x=20, y=20, z=10, kind=S.EXPLOSION, amplitude=1e8)
    a = LA.session(ctx, model)
    a.abc_attach(None)
    assert a.abc_frame_cells() > 0
    a.abc_detach()
    b = LA.session(ctx, model)
    fa, fb = run_fields(a, src, 120), run_fields(b, src, 120)
    assert all(np.array_equal(x, y) for x, y in zip(fa, fb))
    # Zero output layer: the correction is exactly zero.
    c = LA.session(ctx, model)
    c.abc_attach(LA.init_params(7))
    d = LA.session(ctx, model)
    d.abc_attach(None)
    fc, fd = run_fields(c, src, 120), run_fields(d, src, 120)
    assert all(np.array_equal(x, y) for x, y in zip(fc, fd))
    assert not all(np.array_equal(x, y) for x, y in zip(fa, fd))   # the frame differs from CPML
    for g in (a, b, c, d):
        g.release()
    with pytest.raises(S.InvalidArgument):
        e = S.GPUSimData3D(W.imported_context(40, 40, 40, max_vp=4500.0, pml=10))
        e.abc_attach(None)


def test_training_lowers_loss(gpu):
    tr = LA.Trainer(400_000, seed=5)
    n = LA.generate(tr, runs=2, size=(32, 32, 32), steps=150, every=5, start=20, cells=2000, seed=11)
    assert n > 10_000 and tr.count == n
    p0 = LA.init_params(2)
    before = tr.evaluate(p0)
    p, losses = tr.train(p0, epochs=6, batch=4096, lr=2e-3)
    after = tr.evaluate(p)
    print(f"samples {n}, loss {before:.4e} -> {after:.4e}, epochs {losses}")
    assert np.all(np.isfinite(losses))
    assert after < 0.7 * before
    assert losses[-1] < losses[0]


def test_trained_frame_absorbs(gpu):
    """Reflected energy inside a 48^3 layered model (free top) against a 40-cell CPML run."""
    params = LA.load_weights()
    model = LA.layered_random_model(48, 48, 48, seed=4242)
    src = S.default_source(x=24, y=24, z=12, kind=S.EXPLOSION, amplitude=1e8)
    e = LA.reflection(params, model, src, steps=360, every=20)
    print("reflected energy (relative squared error vs 40-cell CPML):", e)
    assert e["learned"] < 0.25 * e["none"]


def test_learned_frame_is_stable_over_long_runs(gpu):
    """900 steps at 64^3: the dissipative correction keeps the learned frame bounded (unrestricted
    one-step corrections grew without bound here) and still well below truncation."""
    params = LA.load_weights()
    model = LA.layered_random_model(64, 64, 64, seed=77)
    src = S.default_source(x=32, y=32, z=16, kind=S.EXPLOSION, amplitude=1e8)
    e = LA.reflection(params, model, src, steps=900, every=100, cases=("learned", "none"))
    print("reflected energy over 900 steps:", e)
    assert np.isfinite(e["learned"]) and e["learned"] < 0.25 * e["none"]
