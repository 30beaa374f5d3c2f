"""Learned 8-cell absorbing boundary (no reference counterpart): data, training and evaluation.

The frame (C ABI ``seis3d_abc_*``): a pml = 8 context whose 8 CPML cells per absorbing face
take the plain update; the outer 4 layers, whose O(8) stencils reach the zero ghost cells, take
a correction from a pointwise MLP (96 -> 32 -> 32 -> 9, BF16 weights on MFMA, FP32
accumulation) after every step. Plain 8-cell CPML (detached) is the fallback.

Training is teacher-forced on wide-CPML runs. A *wide* session models the physical volume grown
by 8 cells on every absorbing face, inside a 40-cell CPML; its margin is the free-space
continuation of the waves that the frame should reproduce. At sampled steps the wide state is
copied into a *thin* session (the physical volume with the frame, plain update only), the thin
session takes one step, the wide one takes the same step, and each sampled frame cell adds the
pair (normalised features, normalised wide - thin difference) to the GPU trainer. The trainer's
forward, backward and Adam run on this library's kernels.

Evaluation: the wavefield inside the physical volume against a 40-cell CPML run of the same
volume (the reflection-free reference); the relative squared error over sampled steps measures
the reflected energy. Compared: the learned frame, the frame without correction (truncation),
plain 8-cell CPML and plain 20-cell CPML.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import seis3d as S
from . import workloads as W

N_PARAMS = 4457
SHAPES = (("W1", (96, 32)), ("b1", (32,)), ("W2", (32, 32)), ("b2", (32,)), ("W3", (32, 9)), ("b3", (9,)))
# Shipped weights, trained on one MI355X by this module's own trainer:
#   python -m cmakepro_cuda_seissim_v2_amd.learned_abc train --runs 200 --size 128 --epochs 40 \
#       --out cmakepro_cuda_seissim_v2_amd/data/abc_weights.npy
# (10.8 M teacher-forced samples from 200 seeded layered 128^3 runs; loss 7.1e-5 -> 1.75e-7).
WEIGHTS_PATH = Path(__file__).resolve().parent / "data" / "abc_weights.npy"
MARGIN = 8
WIDE_PML = 40


def unpack(p: np.ndarray) -> dict:
    out, k = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[k:k + n], np.float32).reshape(shape)
        k += n
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], np.float32).reshape(-1) for name, _ in SHAPES]).astype(np.float32)


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised hidden layers and a zero output layer (the untrained frame adds nothing)."""
    rng = np.random.default_rng(seed)
    d = {"W1": rng.standard_normal((96, 32)) * np.sqrt(2 / 96), "b1": np.zeros(32),
         "W2": rng.standard_normal((32, 32)) * np.sqrt(2 / 32), "b2": np.zeros(32),
         "W3": np.zeros((32, 9)), "b3": np.zeros(9)}
    return pack(d)


def load_weights(path: Path = WEIGHTS_PATH) -> np.ndarray:
    p = np.load(path, allow_pickle=False).astype(np.float32)
    if p.size != N_PARAMS:
        raise ValueError(f"{path}: expected {N_PARAMS} parameters, found {p.size}")
    return p


class Trainer:
    """GPU trainer of the frame MLP (seis3d_abc_trainer_*)."""

    def __init__(self, capacity: int, seed: int = 0):
        self._h = C.c_void_p()
        S._check(S.library().seis3d_abc_trainer_create(int(capacity), int(seed), C.byref(self._h)))

    def close(self) -> None:
        if self._h.value:
            S.library().seis3d_abc_trainer_destroy(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def count(self) -> int:
        return int(S.library().seis3d_abc_trainer_count(self._h))

    def add(self, thin: S.GPUSimData3D, wide: S.GPUSimData3D, offset, max_cells: int, min_rel_scale: float) -> int:
        got = C.c_int()
        S._check(S.library().seis3d_abc_trainer_add(self._h, thin._h, wide._h, int(offset[0]), int(offset[1]),
                                                    int(offset[2]), int(max_cells), float(min_rel_scale),
                                                    C.byref(got)))
        return got.value

    def train(self, params: np.ndarray, epochs: int, batch: int = 8192, lr: float = 2e-3):
        p = np.ascontiguousarray(params, np.float32).copy()
        losses = np.zeros(epochs, np.float32)
        S._check(S.library().seis3d_abc_train(self._h, S._fptr(p), int(epochs), int(batch), float(lr),
                                              S._fptr(losses)))
        return p, losses

    def evaluate(self, params: np.ndarray) -> float:
        p = np.ascontiguousarray(params, np.float32)
        out = C.c_double()
        S._check(S.library().seis3d_abc_evaluate(self._h, S._fptr(p), C.byref(out)))
        return out.value


def mlp(params: np.ndarray, X: np.ndarray, T: np.ndarray):
    """The training kernels on host data: (Y, H1, H2, grad, loss)."""
    p = np.ascontiguousarray(params, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    T = np.ascontiguousarray(T, np.float32)
    n = X.shape[0]
    Y = np.zeros((n, 9), np.float32)
    H1 = np.zeros((n, 32), np.float32)
    H2 = np.zeros((n, 32), np.float32)
    g = np.zeros(N_PARAMS, np.float32)
    loss = C.c_double()
    S._check(S.library().seis3d_abc_mlp(S._fptr(p), S._fptr(X), S._fptr(T), n, S._fptr(Y), S._fptr(H1), S._fptr(H2),
                                        S._fptr(g), C.byref(loss)))
    return Y, H1, H2, g, loss.value


# ---- models and contexts ----------------------------------------------------------------------------

def layered_random_model(nx: int, ny: int, nz: int, seed: int):
    """Five seeded layers (vp 1800..4500 m/s) times (1 + 0.05 g), g a Gaussian random field of
    correlation length 8 cells; vs = vp / 1.73, Gardner density."""
    vp, _, _ = W.layered_model(nx, ny, nz, seed=seed)
    g = W.gaussian_random_field((nz, ny, nx), 8.0, seed + 7)
    vp = (vp * (1.0 + 0.05 * g)).astype(np.float32)
    return vp, (vp / np.float32(1.73)).astype(np.float32), W.gardner_density(vp)


def contexts(nx: int, ny: int, nz: int, max_vp: float, free_surface: bool, frequency: float):
    """(thin pml 8, wide margin + 40-cell CPML, wide offset of the thin grid, margin origin)."""
    thin = W.imported_context(nx, ny, nz, max_vp=max_vp, pml=8, free_surface=free_surface, frequency=frequency)
    mz = 0 if free_surface else MARGIN
    wide = W.imported_context(nx + 2 * MARGIN, ny + 2 * MARGIN, nz + MARGIN + mz, max_vp=max_vp, pml=WIDE_PML,
                              free_surface=free_surface, frequency=frequency)
    # thin padded cell p maps to wide padded cell p + offset (same physical point).
    off = (wide.ox + MARGIN - thin.ox, wide.oy + MARGIN - thin.oy, wide.oz + mz - thin.oz)
    return thin, wide, off, (MARGIN, MARGIN, mz)


def grow(model, margin, free_surface: bool):
    """Clamp-extend a [nz][ny][nx] model by `margin` cells on every absorbing face."""
    mz = 0 if free_surface else margin
    return tuple(np.pad(a, ((mz, margin), (margin, margin), (margin, margin)), mode="edge").astype(np.float32)
                 for a in model)


def session(ctx, model):
    g = S.GPUSimData3D(ctx)
    g.upload_model(*model)
    return g


def shifted(src, origin):
    return S.default_source(x=src.x + origin[0], y=src.y + origin[1], z=src.z + origin[2], kind=src.kind,
                            amplitude=src.amplitude, direction=tuple(src.direction))


def generate(trainer: Trainer, runs: int, size=(48, 48, 48), steps: int = 240, every: int = 4, start: int = 20,
             cells: int = 3000, min_rel_scale: float = 1e-3, seed: int = 1, free_surface: bool = True,
             frequency: float | None = None, log=None) -> int:
    """Add teacher-forced samples of `runs` seeded wide-CPML runs to the trainer (source frequency
    `frequency`, or drawn per run from 12..30 Hz)."""
    nx, ny, nz = size
    rng = np.random.default_rng(seed)
    total = 0
    for r in range(runs):
        model = layered_random_model(nx, ny, nz, seed + 1000 * r)
        f0 = frequency if frequency is not None else float(rng.uniform(12.0, 30.0))
        thin_c, wide_c, off, origin = contexts(nx, ny, nz, float(model[0].max()), free_surface, f0)
        wide = session(wide_c, grow(model, MARGIN, free_surface))
        thin = session(thin_c, model)
        thin.abc_attach(None)
        kind = S.EXPLOSION if rng.random() < 0.5 else S.FORCE
        d = rng.standard_normal(3)
        d = d / np.linalg.norm(d)
        src = S.default_source(x=int(rng.integers(nx // 4, 3 * nx // 4)), y=int(rng.integers(ny // 4, 3 * ny // 4)),
                               z=int(rng.integers(2, nz // 2)), kind=kind, amplitude=1e8, direction=tuple(d))
        wsrc = shifted(src, origin)
        added = 0
        for n in range(steps):
            if n >= start and (n - start) % every == 0:
                thin.copy_wavefield_box(wide, *off)
                thin.step(src, n)
                wide.step(wsrc, n)
                added += trainer.add(thin, wide, off, cells, min_rel_scale)
            else:
                wide.step(wsrc, n)
        thin.release()
        wide.release()
        total += added
        if log:
            log(f"run {r}: {added} samples (total {trainer.count})")
    return total


# ---- evaluation ---------------------------------------------------------------------------------------

def reflection(params, model, src, steps: int, every: int, free_surface: bool = True, frequency: float = 25.0,
               cases=("learned", "none", "cpml8", "cpml20"), log=None, profile=None) -> dict:
    """Relative squared velocity error inside the physical volume against a 40-cell CPML run,
    summed over every `every`-th step: the reflected energy of each boundary ("profile": 8-cell
    CPML with the learned profile, `profile` or the shipped one).  `log(message)`, if given, is
    called after each sampled step."""
    nz, ny, nx = model[0].shape
    vmax = float(model[0].max())
    ref_c = W.imported_context(nx, ny, nz, max_vp=vmax, pml=WIDE_PML, free_surface=free_surface, frequency=frequency)
    ref = session(ref_c, model)
    runs = {}
    for case in cases:
        pml = 20 if case == "cpml20" else 8
        c = W.imported_context(nx, ny, nz, max_vp=vmax, pml=pml, free_surface=free_surface, frequency=frequency)
        g = session(c, model)
        if case == "learned":
            g.abc_attach(params)
        elif case == "none":
            g.abc_attach(None)
        elif case == "profile":
            g.set_cpml_profile(profile if profile is not None else load_profile())
        runs[case] = (c, g)
    err = {case: 0.0 for case in cases}
    norm = 0.0
    for n in range(steps):
        ref.step(src, n)
        for case in cases:
            runs[case][1].step(src, n)
        if (n + 1) % every:
            continue
        for f in (0, 1, 2):
            r = ref.download_field(f)[ref_c.oz:ref_c.oz + nz, ref_c.oy:ref_c.oy + ny, ref_c.ox:ref_c.ox + nx]
            norm += float(np.sum(r.astype(np.float64) ** 2))
            for case in cases:
                c, g = runs[case]
                v = g.download_field(f)[c.oz:c.oz + nz, c.oy:c.oy + ny, c.ox:c.ox + nx]
                err[case] += float(np.sum((v.astype(np.float64) - r) ** 2))
        if log:
            log(f"reflection: step {n + 1}/{steps} " + ", ".join(f"{c} {err[c] / norm:.3e}" for c in cases))
    for c, g in runs.values():
        g.release()
    ref.release()
    return {case: err[case] / norm for case in cases}


# ---- learned CPML profile ---------------------------------------------------------------------------
#
# The second learned boundary: 8-cell CPML whose profile shapes (seis3d_set_cpml_profile: sigma scale,
# sigma/kappa power, kappa max, alpha scale, alpha power) are fitted to wide-CPML runs.  The step
# kernels are the plain CPML ones, so the per-step cost is exactly that of 8-cell CPML with the
# reference profile.  Fitted by Nelder-Mead on the mean log10 reflected energy of seeded training
# cases (GPU rollouts against stored 40-cell CPML references); judged on held-out seeds.
# Shipped profile:
#   python -m cmakepro_cuda_seissim_v2_amd.learned_abc profile --out cmakepro_cuda_seissim_v2_amd/data/cpml_profile.json
PROFILE_PATH = Path(__file__).resolve().parent / "data" / "cpml_profile.json"
PROFILE_TRAIN_SEEDS = tuple(range(101, 109))
PROFILE_HELD_OUT_SEEDS = tuple(range(901, 905))


def load_profile(path: Path = PROFILE_PATH) -> tuple:
    import json
    d = json.loads(Path(path).read_text())
    p = tuple(float(v) for v in d["profile"])
    if len(p) != S.PROFILE_PARAMS:
        raise ValueError(f"{path}: a CPML profile has {S.PROFILE_PARAMS} parameters")
    return p


def profile_case(seed: int, size=(64, 64, 64), steps: int = 360, every: int = 8, free_surface: bool = True) -> dict:
    """A seeded case: layered random model, a source of random kind, position, direction and
    frequency (12..30 Hz), and the velocities of a 40-cell CPML run of the same volume inside the
    physical volume at every `every`-th step (host, float32)."""
    nx, ny, nz = size
    rng = np.random.default_rng(seed)
    model = layered_random_model(nx, ny, nz, seed)
    f0 = float(rng.uniform(12.0, 30.0))
    kind = S.EXPLOSION if rng.random() < 0.5 else S.FORCE
    d = rng.standard_normal(3)
    d = d / np.linalg.norm(d)
    src = S.default_source(x=int(rng.integers(nx // 4, 3 * nx // 4)), y=int(rng.integers(ny // 4, 3 * ny // 4)),
                           z=int(rng.integers(2, nz // 2)), kind=kind, amplitude=1e8, direction=tuple(d))
    vmax = float(model[0].max())
    ref_c = W.imported_context(nx, ny, nz, max_vp=vmax, pml=WIDE_PML, free_surface=free_surface, frequency=f0)
    thin_c = W.imported_context(nx, ny, nz, max_vp=vmax, pml=8, free_surface=free_surface, frequency=f0)
    ref = session(ref_c, model)
    snaps = []
    norm = 0.0
    for n in range(steps):
        ref.step(src, n)
        if (n + 1) % every == 0:
            frame = [ref.download_field(f)[ref_c.oz:ref_c.oz + nz, ref_c.oy:ref_c.oy + ny, ref_c.ox:ref_c.ox + nx]
                     .astype(np.float32) for f in (0, 1, 2)]
            norm += sum(float(np.sum(v.astype(np.float64) ** 2)) for v in frame)
            snaps.append(frame)
    ref.release()
    return {"seed": seed, "model": model, "src": src, "ctx": thin_c, "steps": steps, "every": every,
            "snaps": snaps, "norm": norm, "frequency": f0}


def profile_reflection(profile, case: dict) -> float:
    """Reflected energy (relative squared velocity error in the physical volume, summed over the
    case's sampled steps) of 8-cell CPML with `profile` (None: the reference profile)."""
    c = case["ctx"]
    nz, ny, nx = case["model"][0].shape
    g = session(c, case["model"])
    if profile is not None:
        g.set_cpml_profile(profile)
    err = 0.0
    k = 0
    for n in range(case["steps"]):
        g.step(case["src"], n)
        if (n + 1) % case["every"] == 0:
            for f in (0, 1, 2):
                v = g.download_field(f)[c.oz:c.oz + nz, c.oy:c.oy + ny, c.ox:c.ox + nx]
                err += float(np.sum((v.astype(np.float64) - case["snaps"][k][f]) ** 2))
            k += 1
    g.release()
    return err / case["norm"]


def train_profile(cases, start=S.REFERENCE_PROFILE, max_evals: int = 160, log=None):
    """Nelder-Mead over the 5 profile parameters, minimising the mean log10 reflected energy over
    `cases`; parameters outside the accepted ranges score +inf.  Returns (profile, history)."""
    from scipy.optimize import minimize

    lo = np.array([-5.0, 1.0, 1.0, 0.0, 0.1])
    hi = np.array([5.0, 8.0, 20.0, 10.0, 8.0])
    history = []

    def objective(x):
        if np.any(x < lo) or np.any(x > hi):
            return np.inf
        e = [profile_reflection(tuple(x), c) for c in cases]
        score = float(np.mean(np.log10(e)))
        history.append((tuple(float(v) for v in x), score))
        if log:
            log(f"profile eval {len(history)}: {np.round(x, 4).tolist()} mean log10 {score:.4f}")
        return score

    x0 = np.asarray(start, dtype=np.float64)
    simplex = [x0] + [x0 + np.eye(5)[i] * step for i, step in enumerate((0.3, 0.6, 0.8, 0.4, 0.6))]
    res = minimize(objective, x0, method="Nelder-Mead",
                   options={"maxfev": max_evals, "initial_simplex": np.array(simplex), "xatol": 1e-3, "fatol": 1e-3})
    return tuple(float(v) for v in res.x), history


def main(argv=None) -> int:
    """``python -m cmakepro_cuda_seissim_v2_amd.learned_abc train|reflect ...``"""
    import argparse
    import json
    import sys
    import time

    ap = argparse.ArgumentParser(prog="learned_abc")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train", help="generate teacher-forced samples from wide-CPML runs and train on the GPU")
    t.add_argument("--runs", type=int, default=200)
    t.add_argument("--size", type=int, default=128)
    t.add_argument("--steps", type=int, default=0, help="steps per run (0: 2.2 x size)")
    t.add_argument("--every", type=int, default=4)
    t.add_argument("--cells", type=int, default=1000)
    t.add_argument("--epochs", type=int, default=12)
    t.add_argument("--batch", type=int, default=16384)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--out", default=str(WEIGHTS_PATH))
    q = sub.add_parser("profile", help="fit the learned CPML profile on GPU rollouts against 40-cell CPML runs")
    q.add_argument("--size", type=int, default=64)
    q.add_argument("--steps", type=int, default=360)
    q.add_argument("--every", type=int, default=8)
    q.add_argument("--evals", type=int, default=160)
    q.add_argument("--out", default=str(PROFILE_PATH))
    j = sub.add_parser("reflect", help="reflected energy of the learned frame against a 40-cell CPML run")
    j.add_argument("--size", default="300x300x400")
    j.add_argument("--steps", type=int, default=1500)
    j.add_argument("--every", type=int, default=100)
    j.add_argument("--weights", default=str(WEIGHTS_PATH))
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if a.cmd == "train":
        steps = a.steps or int(2.2 * a.size)
        per_run = ((steps - 20) // a.every + 1) * a.cells
        tr = Trainer(min(per_run * a.runs, 1 << 26), seed=a.seed)
        t0 = time.perf_counter()
        n = generate(tr, a.runs, size=(a.size,) * 3, steps=steps, every=a.every, cells=a.cells, seed=a.seed,
                     log=lambda m: log(m) if (int(m.split()[1].rstrip(":")) % 20 == 0) else None)
        t_gen = time.perf_counter() - t0
        p0 = init_params(a.seed)
        before = tr.evaluate(p0)
        t0 = time.perf_counter()
        p, losses = tr.train(p0, a.epochs, a.batch, a.lr)
        t_train = time.perf_counter() - t0
        after = tr.evaluate(p)
        Path(a.out).parent.mkdir(parents=True, exist_ok=True)
        np.save(a.out, p.astype(np.float32), allow_pickle=False)
        print(json.dumps({"samples": n, "runs": a.runs, "size": a.size, "steps": steps, "generate_s": round(t_gen, 2),
                          "train_s": round(t_train, 2), "loss_before": before, "loss_after": after,
                          "epoch_loss": [float(x) for x in losses], "out": a.out}), flush=True)
        return 0
    if a.cmd == "profile":
        size = (a.size,) * 3
        t0 = time.perf_counter()
        train, held = [], []
        for group, seeds in ((train, PROFILE_TRAIN_SEEDS), (held, PROFILE_HELD_OUT_SEEDS)):
            for sd in seeds:
                group.append(profile_case(sd, size, a.steps, a.every))
                log(f"profile reference case {sd}: {time.perf_counter() - t0:.1f} s")
        t_ref = time.perf_counter() - t0
        t0 = time.perf_counter()
        best, history = train_profile(train, max_evals=a.evals, log=log)
        t_fit = time.perf_counter() - t0
        rows = {}
        for name, group in (("train", train), ("held_out", held)):
            rows[name] = [{"seed": c["seed"], "frequency_hz": round(c["frequency"], 2),
                           "reference_profile": profile_reflection(None, c), "learned_profile": profile_reflection(best, c)}
                          for c in group]
        out = {"profile": list(best), "reference_profile": list(S.REFERENCE_PROFILE), "size": a.size, "steps": a.steps,
               "every": a.every, "evaluations": len(history), "references_s": round(t_ref, 1), "fit_s": round(t_fit, 1),
               "cases": rows,
               "held_out_gain": [r["reference_profile"] / r["learned_profile"] for r in rows["held_out"]]}
        Path(a.out).parent.mkdir(parents=True, exist_ok=True)
        Path(a.out).write_text(json.dumps(out, indent=1) + "\n")
        print(json.dumps(out), flush=True)
        return 0
    nx, ny, nz = (int(v) for v in a.size.split("x"))
    params = load_weights(Path(a.weights))
    model = layered_random_model(nx, ny, nz, seed=20261203)
    src = S.default_source(x=nx // 2, y=ny // 2, z=max(2, nz // 4), kind=S.EXPLOSION, amplitude=1e8)
    t0 = time.perf_counter()
    e = reflection(params, model, src, a.steps, a.every, log=log)
    print(json.dumps({"grid": [nx, ny, nz], "steps": a.steps, "every": a.every, "reflected_energy": e,
                      "gain_vs_none": e["none"] / e["learned"], "gain_vs_cpml8": e["cpml8"] / e["learned"],
                      "seconds": round(time.perf_counter() - t0, 1)}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
