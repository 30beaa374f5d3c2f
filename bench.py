#!/usr/bin/env python3
"""Benchmark of the fused 3D O(8) velocity-stress step on one MI355X.

Workload (one step = one leapfrog time step of the hot path): 300 x 50 x 300 grid, dx = 10 m,
free surface on top and 20-cell CPML on the other five faces, five seeded horizontal layers
(vp 1800..4500 m/s, vs = vp/1.73, Gardner density), 15 Hz Ricker vertical force, the default
150-receiver surface spread sampled every 2 steps into the device gather.  dt is the
reference's automatic CFL step, 847 us: the requested 1 ms violates the O(8) 3D CFL bound and
the reference rejects it.

Prints ONE JSON line: Gcell-updates/s over the timed steps (inputs resident in HBM), the HBM
roofline of the dominant kernel measured with HIP events on the session stream, and the CPU
oracle timed on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "Gcell-updates/s 3D O(8) VS-FD + %HBM roofline; batched FWI-grad shots/s vs plain"
FAMILY_NAMES = {0: "direct", 4: "lds-dma-2.5D"}
FAMILY_KERNEL = {0: "direct", 4: "dma"}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--warmup", type=int, default=50)
    p.add_argument("--family", type=int, default=-1, help="kernel family (-1: library default)")
    p.add_argument("--tile-y", type=int, default=0)
    p.add_argument("--z-chunk", type=int, default=-1,
                   help="planes per workgroup (0: balanced one-round chunks with --tile-y 8; default: autotuned)")
    p.add_argument("--kernel-steps", type=int, default=200, help="steps timed per kernel with HIP events")
    p.add_argument("--cpu-steps", type=int, default=0, help="oracle steps for the CPU baseline (0 = auto)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--grid", default="300x50x300")
    p.add_argument("--workload", choices=["step", "batch", "fwi", "2d", "abc"], default="step",
                   help="step: the fused 3D step (default line); batch: 16 uneven shots, graph vs plain loop")
    p.add_argument("--shots", type=int, default=0, help="shots of the batch (16) or fwi (8) workload")
    p.add_argument("--checkpoint", type=int, default=50, help="FWI checkpoint interval (steps)")
    p.add_argument("--fd", action="store_true",
                   help="FWI: central-difference checks along the vp, vs and rho parts of the gradient")
    p.add_argument("--no-copy-ceiling", action="store_true",
                   help="skip the stream-kernel ceiling measurement (profiled runs)")
    p.add_argument("--copy-ceiling", action="store_true",
                   help="instead of the bench, print the measured stream-kernel ceilings (GB/s)")
    p.add_argument("--no-extra", action="store_true",
                   help="default line only: skip its grid_200, fwi and batch objects (profiled runs)")
    p.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                   help="instead of the bench, run 5 launches of the library's copy kernel of MIB MiB (counter calibration)")
    return p.parse_args()


def copy_probe(mib: int) -> None:
    """Five launches of this library's copy kernel over arrays of a known size, to calibrate
    rocprofv3's FETCH_SIZE / WRITE_SIZE counters against a kernel whose HBM traffic is known."""
    from cmakepro_cuda_seissim_v2_amd import seis3d as S
    S.bandwidth_probe("copy", mib, 5)
    print(json.dumps({"copy_probe_mib": mib, "bytes_read_per_copy": mib << 20, "bytes_written_per_copy": mib << 20}))


def measured_ceilings() -> dict:
    """The measured HBM ceilings: this library's own stream kernels (seis3d_bandwidth_probe), GB/s
    of bytes moved (read + written).  read/write/copy use 16-byte lanes; copy_lds_dma stages
    through LDS with global_load_lds like the step kernels; mix10_3 and mix12_6 stream the
    velocity and stress half-steps' numbers of arrays elementwise."""
    from cmakepro_cuda_seissim_v2_amd import seis3d as S
    out = {}
    for kind, mib in (("read", 1024), ("write", 1024), ("copy", 1024), ("copy_lds_dma", 1024), ("mix10_3", 256),
                      ("mix12_6", 256)):
        try:
            out[kind] = round(S.bandwidth_probe(kind, mib, 10)["gbs"], 1)
        except Exception as e:  # the line still carries the paper peak
            print(f"bench: probe {kind} unavailable ({e})", file=sys.stderr, flush=True)
    print(f"bench: measured ceilings GB/s {out}", file=sys.stderr, flush=True)
    return out


def profiled_traffic(kernel: str, grid: str, tiles):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary of this workload
    (profiles/<family>_kernels_<grid>.md, line 'traffic_bytes_per_launch: tiles=<tile_y>x<z_chunk>
    name=bytes ...'), when it was measured with the same kernel shape."""
    shape = f"{tiles[1]}x{tiles[2]}"
    for path in sorted((ROOT / "profiles").glob(f"*_{grid}.md")):
        for line in path.read_text().splitlines():
            if line.startswith("traffic_bytes_per_launch:"):
                items = dict(item.partition("=")[::2] for item in line.split(":", 1)[1].split())
                if items.get("tiles") == shape and kernel in items:
                    return int(items[kernel]), path.name
    return None, None


def batch_run(steps: int, n_shots: int = 16) -> dict:
    """16 shots on uneven models (128^3 .. 256x128x256), `steps` steps, 100 receivers each, stride 2.

    plain: the reference's shot loop, one shot after another, runStep3D per step and
    sampleReceivers3D (launch + copy + stream sync) every stride (src/Seis3D_Sim.cpp:120-153).
    batched: all steps of every shot captured in one HIP graph, one launch, device-side gathers, in
    each of the three batch modes (a stream per shot; shared launches over groups of up to four
    shots; one persistent work-queue launch per half-step for all shots).  Model upload, capture
    and instantiation are setup, outside every timing.
    """
    import numpy as np
    from cmakepro_cuda_seissim_v2_amd import seis3d as S
    from cmakepro_cuda_seissim_v2_amd import workloads as W

    stride = 2
    shots = W.batch_shots(n_shots)
    sessions = []
    for ctx, model, src, rec in shots:
        g = S.GPUSimData3D(ctx)
        g.upload_model(*model)
        g.set_receivers(rec)
        sessions.append(g)
    tuned = {}
    for (ctx, model, src, rec), g in zip(shots, sessions):
        key = (ctx.gx, ctx.gy, ctx.gz)
        tuned[key] = g.autotune(src, steps=3)
    n_upd = sum((c.gx - 8) * (c.gy - 8) * (c.gz - 8) for c, _, _, _ in shots)
    print(f"bench: batch of {len(shots)} shots, {n_upd / 1e6:.1f} M updated cells", file=sys.stderr, flush=True)

    # Plain per-shot loop.
    for g in sessions:
        g.reset()
        g.synchronize()
    t0 = time.perf_counter()
    plain = []
    for (ctx, model, src, rec), g in zip(shots, sessions):
        samples = []
        for n in range(steps):
            g.step(src, n)
            if (n + 1) % stride == 0:
                samples.append(g.sample_receivers())
        g.synchronize()
        plain.append(np.stack(samples))
    t_plain = time.perf_counter() - t0
    print(f"bench: plain loop {t_plain:.3f} s", file=sys.stderr, flush=True)

    times, same = {}, True
    for mode in ("shot_streams", "shared_launches", "work_queue"):
        for g in sessions:
            g.reset()
            g.gather_end()
            g.gather_begin(steps // stride, stride)
        batch = S.Batch(sessions, [s[2] for s in shots], steps, mode=mode)
        for g in sessions:
            g.synchronize()
        t0 = time.perf_counter()
        batch.launch()
        batch.synchronize()
        times[mode] = time.perf_counter() - t0
        ok = all(np.array_equal(g.gather_read(), p) for g, p in zip(sessions, plain))
        same = same and ok
        print(f"bench: batched graph ({mode}) {times[mode]:.3f} s, gathers identical to plain: {ok}",
              file=sys.stderr, flush=True)
        batch.close()
    for g in sessions:
        g.release()
    return {
        "workload": f"{len(shots)} shots, nx,nz in 128..256 step 16, ny 128, 20-cell CPML, free top, "
                    f"{steps} steps, 15 Hz force, 100 receivers stride {stride}",
        "shots": len(shots), "steps": steps,
        "updated_cells_total": n_upd,
        "plain_s": round(t_plain, 3),
        "plain_shots_per_s": round(len(shots) / t_plain, 4),
        "batched_s": {m: round(t, 3) for m, t in times.items()},
        "batched_shots_per_s": {m: round(len(shots) / t, 4) for m, t in times.items()},
        "speedup_vs_plain": {m: round(t_plain / t, 3) for m, t in times.items()},
        "gathers_bit_identical": bool(same),
        "gcell_updates_per_s_batched": {m: round(n_upd * steps / t / 1e9, 3) for m, t in times.items()},
        "autotuned": {f"{k[0]}x{k[1]}x{k[2]}": v for k, v in tuned.items()},
    }


def batch_bench(args) -> None:
    """The batch workload on its own line (value: the default per-shot-stream mode)."""
    steps = args.steps
    b = batch_run(steps, args.shots or 16)
    t_batch = b["batched_s"]["shot_streams"]
    out = {
        "metric": "shots/s, 16-shot batched 3D O(8) forward (one hipGraph launch) vs the plain per-shot loop",
        "value": b["batched_shots_per_s"]["shot_streams"],
        "unit": "shots/s",
        "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round(t_batch / steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded gradient + Gaussian random-field models)",
        "config": b,
    }
    print(json.dumps(out), flush=True)


def fwi_misfit(S, ctx, model, shots, observed, steps, stride):
    """0.5 |gather - observed|^2 over all shots from plain forward runs (FP32 GPU, FP64 sum)."""
    import numpy as np
    total = 0.0
    with S.GPUSimData3D(ctx) as g:
        g.upload_model(*model)
        for (src, rec), obs in zip(shots, observed):
            g.reset()
            g.set_receivers(rec)
            g.gather_begin(steps // stride, stride)
            g.run(src, 0, steps)
            got = g.gather_read()
            g.gather_end()
            total += 0.5 * float(np.sum((got.astype(np.float64) - obs.astype(np.float64)) ** 2))
    return total


def fwi_fd_checks(S, W, ctx, start, gradient, shots, observed, steps, stride, params=(0, 1)) -> list:
    """Central differences of the misfit along `fwi_check_direction` of each parameter in `params`,
    at relative steps 5e-4 and 1e-3, Richardson-extrapolated (fourth order in the step)."""
    import numpy as np
    base = [np.asarray(a, np.float32) for a in start]
    out = []
    for k in params:
        d = W.fwi_check_direction(ctx, base, gradient[1 + k], k)
        ratios = []
        for eps in (5e-4, 1e-3):
            js, moved = [], []
            for sign in (1, -1):
                m = list(base)
                m[k] = (base[k].astype(np.float64) + sign * eps * d).astype(np.float32)
                moved.append(m[k])
                js.append(fwi_misfit(S, ctx, tuple(m), shots, observed, steps, stride))
            pred = float(np.sum(gradient[1 + k].astype(np.float64) * (moved[0].astype(np.float64) - moved[1])))
            ratios.append((js[0] - js[1]) / pred)
        extrap = (4 * ratios[0] - ratios[1]) / 3
        out.append({"direction": ["vp", "vs", "rho"][k] + " part of the gradient, below 20-30 cells",
                    "fd_over_adjoint": [round(r, 7) for r in ratios], "extrapolated": round(extrap, 7),
                    "rel_error": float(f"{abs(extrap - 1):.3e}")})
        print(f"bench: fd direction {out[-1]}", file=sys.stderr, flush=True)
    return out


def fwi_bench(args) -> None:
    """FWI gradient of 8 shots on 200x100x200, free top (1200 steps, checkpoint every 50 steps).

    batched: one seis3d_fwi_gradient request with every shot (shots on their own streams, as many
    at once as fit in HBM). plain: the same library called once per shot, one after another.
    Observed data come from forward runs on the true model (outside both timings).
    """
    import numpy as np
    from cmakepro_cuda_seissim_v2_amd import seis3d as S
    from cmakepro_cuda_seissim_v2_amd import workloads as W

    steps, stride, interval = args.steps, 2, args.checkpoint
    ctx, start, true, shots = W.fwi_workload(args.shots or 8)
    plan = S.fwi_plan(ctx, steps, stride, interval, len(shots[0][1]))
    print(f"bench: fwi {len(shots)} shots {ctx.nx}x{ctx.ny}x{ctx.nz}, {steps} steps, "
          f"{plan['shot_bytes'] / 2**30:.1f} GiB per shot", file=sys.stderr, flush=True)
    observed = []
    with S.GPUSimData3D(ctx) as g:
        g.upload_model(*true)
        for src, rec in shots:
            g.reset()
            g.set_receivers(rec)
            g.gather_begin(steps // stride, stride)
            g.run(src, 0, steps)
            observed.append(g.gather_read())
            g.gather_end()
    print("bench: observed data ready", file=sys.stderr, flush=True)
    req = [(src, rec, obs) for (src, rec), obs in zip(shots, observed)]
    with S.GPUSimData3D(ctx) as g:
        g.upload_model(*start)
        g.synchronize()
        t0 = time.perf_counter()
        batched = g.fwi_gradient(req, steps, stride, interval)
        t_batch = time.perf_counter() - t0
        print(f"bench: batched gradient {t_batch:.3f} s", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        singles = [g.fwi_gradient([r], steps, stride, interval) for r in req]
        t_plain = time.perf_counter() - t0
        print(f"bench: plain per-shot loop {t_plain:.3f} s", file=sys.stderr, flush=True)
    summed = [np.sum([s[k].astype(np.float64) for s in singles], axis=0) for k in (1, 2, 3)]
    agree = max(float(np.linalg.norm(batched[k + 1] - summed[k]) / np.linalg.norm(summed[k])) for k in range(3))
    n_upd = (ctx.gx - 8) * (ctx.gy - 8) * (ctx.gz - 8)
    fd = fwi_fd_checks(S, W, ctx, start, batched, shots, observed, steps, stride, params=(0, 1, 2)) if args.fd else []
    out = {
        "metric": "FWI gradient shots/s (vp, vs, rho; discrete adjoint + checkpoints), batched vs plain per-shot loop",
        "value": round(len(shots) / t_batch, 4),
        "unit": "shots/s",
        "n_gpus": 1, "steps": steps, "warmup": 0,
        "ms_per_step": round(t_batch / steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (smooth gradient background + seeded Gaussian anomaly, +8 % vp, radius 15 cells)",
        "config": {
            "workload": f"FWI gradient, {len(shots)} shots, {ctx.nx}x{ctx.ny}x{ctx.nz}, dx=10 m, "
                        f"{'free top, CPML 20 cells on 5 faces' if ctx.free_surface else 'all-CPML 20 cells'}, "
                        f"{steps} steps, 15 Hz explosions, {len(shots[0][1])} "
                        f"{'surface ' if ctx.free_surface else ''}receivers stride {stride}, "
                        f"checkpoint every {interval} steps",
            "dt_us": round(ctx.dt * 1e6), "n_upd": n_upd, "shot_bytes": plan["shot_bytes"],
            "plain_shots_per_s": round(len(shots) / t_plain, 4),
            "speedup_vs_plain": round(t_plain / t_batch, 3),
            "batched_vs_plain_rel_l2": float(f"{agree:.3e}"),
            "misfit": batched[0],
            "fd_checks": fd,
        },
    }
    print(json.dumps(out), flush=True)


def bench_2d(args) -> None:
    """Config 1: 2D P-SV, 400 x 400 (PML frame included), h=5 m, dt=0.5 ms, 2000 steps, type-1 split PML,
    25 Hz explosive at (200, 200), 100 receivers recorded every step.

    GPU: seis2d_run (a captured 64-step graph replayed, device-side traces), after one untimed run.
    CPU: the FP32 oracle (the reference's kernels restated in C++) on 1 thread and on all host cores.
    Also the reference's historical 2D figure's shape: 1000 x 500, PML 50 (5.02 GNUPs on an
    RTX 4060 Laptop, a context number from before the O(8) weight fix).
    """
    import numpy as np
    from cmakepro_cuda_seissim_v2_amd import seis2d as S2
    from oracle import pyoracle as O
    from oracle import pyoracle2d as O2

    ctx = S2.config1_context()
    steps = args.steps if args.steps != 1000 else ctx.nt
    ctx.nt = max(ctx.nt, steps)
    ctx.wavelet = S2.generateRickerWavelet(ctx.nt, ctx.dt, 25.0, 1.0 / 25.0)
    g = S2.initializeGpuSimulation(S2.GPUSimData(), ctx)
    g.run(ctx, 0, steps)                       # warm-up (also builds the code objects)
    S2.releaseGpuSimulation(g)
    g = S2.initializeGpuSimulation(S2.GPUSimData(), ctx)
    t0 = time.perf_counter()
    vx, vz = g.run(ctx, 0, steps)
    secs = time.perf_counter() - t0
    S2.releaseGpuSimulation(g)
    n_upd = (ctx.NX - 8) * (ctx.NZ - 8)
    print(f"bench: 2d config 1 {steps} steps in {secs:.4f} s", file=sys.stderr, flush=True)

    big = S2.homogeneous_context(1000, 500, 5.0, 5e-4, 2000, 50, 3000.0, 1732.0, 2000.0, src=(500, 250))
    g = S2.initializeGpuSimulation(S2.GPUSimData(), big)
    g.run(big, 0, 200)
    t0 = time.perf_counter()
    g.run(big, 0, 2000)
    big_secs = time.perf_counter() - t0
    S2.releaseGpuSimulation(g)

    cores, cpu_model = cpu_info()
    cpu = {}
    for threads in (1, min(cores, 16)):
        O.set_threads(threads)
        with O2.Oracle2D(ctx, 32) as o:
            o.time_steps(0, 5)
            probe = o.time_steps(5, 20) / 20
            n = max(20, min(steps, int(12.0 / max(probe, 1e-6))))
            cs = o.time_steps(25, n)
        cpu[threads] = (n, cs)
        print(f"bench: cpu oracle {threads} threads: {n} steps in {cs:.3f} s", file=sys.stderr, flush=True)
    t_all = min(cores, 16)
    n_all, s_all = cpu[t_all]
    O.set_threads(t_all)
    triad = O.stream_triad(1 << 26, t_all)
    # The bench line's own error: the timed GPU run's receiver records against the oracles.
    parity = {"steps": steps, "records": "vx, vz of every receiver at every step"}
    for precision in (32, 64):
        with O2.Oracle2D(ctx, precision) as o:
            rx, rz = o.run(0, steps)
        got = np.concatenate([vx.astype(np.float64).ravel(), vz.astype(np.float64).ravel()])
        ref = np.concatenate([rx.ravel(), rz.ravel()])
        err = float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300))
        parity[f"rel_l2_vs_fp{precision}_oracle"] = float(f"{err:.3e}")
        if precision == 32:
            parity["bit_equal_vs_fp32_oracle"] = bool(np.array_equal(got.astype(np.float32), ref.astype(np.float32)))
    parity["tolerance_fp32"], parity["tolerance_fp64"] = 1e-5, 1e-4
    print(f"bench: 2d parity {parity}", file=sys.stderr, flush=True)
    bytes_cell = 136   # type 1: stress 19 floats (4 v + 6 sigma rmw + 3 moduli), velocity 15 (6 sigma + 4 v rmw + rho)
    cpu_gbs = bytes_cell * n_upd * n_all / s_all / 1e9
    value = n_upd * steps / secs / 1e9
    out = {
        "metric": "Gcell-updates/s 2D P-SV O(8) split-PML step (config 1, reference CPU-path workload)",
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": steps, "warmup": steps,
        "ms_per_step": round(secs / steps * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (homogeneous vp 3000, vs 1732, rho 2000; 25 Hz explosive Ricker)",
        "config": {
            "workload": "2D P-SV O(8), 400x400 incl. 20-cell split PML frame, h=5 m, dt=0.5 ms, explosive at "
                        "(200,200), 100 receivers at row 30 every step",
            "n_upd": n_upd, "grid_cells": ctx.NX * ctx.NZ,
            "whole_grid_gnups": round(ctx.NX * ctx.NZ * steps / secs / 1e9, 4),
            "step_gbs_algorithmic": round(bytes_cell * n_upd * steps / secs / 1e9, 1),
            "grid_1000x500_pml50_gnups": round(1000 * 500 * 2000 / big_secs / 1e9, 4),
            "grid_1000x500_reference_rtx4060_gnups": 5.02,
        },
        "roofline": {
            "bound": "hbm",
            "note": "launch/latency-bound: the 13 MB working set stays in L2/MALL between steps",
            "achieved": round(bytes_cell * n_upd * steps / secs / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(bytes_cell * n_upd * steps / secs / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
        },
        "parity": parity,
        "cpu_baseline": {
            "value": round(n_upd * n_all / s_all / 1e9, 5), "unit": "Gcell-updates/s", "cores": t_all, "kind": "port",
            "one_thread_value": round(n_upd * cpu[1][0] / cpu[1][1] / 1e9, 5),
            "stream_triad_gbs": round(triad, 1),
            "algorithmic_gbs": round(cpu_gbs, 2),
            "frac_of_stream_triad": round(cpu_gbs / triad, 3) if triad > 0 else None,
            "sample": f"FP32 oracle (the reference's type-1 kernels restated in C++, OpenMP over rows) on config 1, "
                      f"{n_all} steps in {s_all:.2f} s on {t_all} threads; 1 thread: "
                      f"{n_upd * cpu[1][0] / cpu[1][1] / 1e9:.5f} Gcell/s; host STREAM triad "
                      f"(3 x 512 MiB doubles, {t_all} threads); host CPU: {cpu_model}",
        },
    }
    print(json.dumps(out), flush=True)


def bench_abc(args) -> None:
    """Learned 8-cell absorbing boundaries on 300 x 300 x 400 (36 M nodes), 1500 steps, free top,
    layered-plus-random-field model: reflected energy against a 40-cell CPML run (the reflection-free
    reference) for the learned frame (MLP weights shipped in the package), the frame without its
    network, plain 8-cell CPML, 8-cell CPML with the learned profile (shipped in the package) and plain
    20-cell CPML (the fallback); and the time per step of each boundary."""
    from cmakepro_cuda_seissim_v2_amd import learned_abc as LA
    from cmakepro_cuda_seissim_v2_amd import seis3d as S
    from cmakepro_cuda_seissim_v2_amd import workloads as W

    nx, ny, nz = (int(v) for v in (args.grid if args.grid != "300x50x300" else "300x300x400").split("x"))
    steps = args.steps if args.steps != 1000 else 1500
    params = LA.load_weights()
    model = LA.layered_random_model(nx, ny, nz, seed=20261203)
    src = S.default_source(x=nx // 2, y=ny // 2, z=max(2, nz // 4), kind=S.EXPLOSION, amplitude=1e8)
    vmax = float(model[0].max())
    times = {}
    profile = LA.load_profile()
    for case, pml in (("learned", 8), ("cpml8", 8), ("profile", 8), ("cpml20", 20), ("cpml40", 40)):
        ctx = W.imported_context(nx, ny, nz, max_vp=vmax, pml=pml, free_surface=True, frequency=25.0)
        g = LA.session(ctx, model)
        if case == "learned":
            g.abc_attach(params)
        elif case == "profile":
            g.set_cpml_profile(profile)
        g.run(src, 0, 20)
        g.synchronize()
        t0 = time.perf_counter()
        g.run(src, 20, 200)
        g.synchronize()
        times[case] = (time.perf_counter() - t0) / 200 * 1e3
        g.release()
        print(f"bench: abc {case} {times[case]:.3f} ms/step", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    e = LA.reflection(params, model, src, steps, every=100, cases=("learned", "none", "cpml8", "profile", "cpml20"),
                      log=lambda m: print(f"bench: abc {m}", file=sys.stderr, flush=True), profile=profile)
    secs = time.perf_counter() - t0
    print(f"bench: abc reflected energy {e} ({secs:.1f} s)", file=sys.stderr, flush=True)
    out = {
        "metric": "reflected energy of the learned 8-cell boundary vs a 40-cell CPML run (relative squared velocity error)",
        "value": float(f"{e['learned']:.4e}"),
        "unit": "relative energy",
        "n_gpus": 1, "steps": steps, "warmup": 20,
        "ms_per_step": round(times["learned"], 4),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32 wavefields, bf16 MLP",
        "data": "synthetic (5 seeded layers x (1 + 0.05 Gaussian random field); 25 Hz explosion)",
        "config": {
            "workload": f"3D O(8), {nx}x{ny}x{nz}, dx=10 m, free top, {steps} steps, sampled every 100 steps",
            "reflected_energy": e,
            "gain_vs_frame_without_network": e["none"] / e["learned"],
            "learned_profile": list(profile),
            "learned_profile_gain_vs_cpml8": e["cpml8"] / e["profile"],
            "ms_per_step": {k: round(v, 4) for k, v in times.items()},
        },
    }
    print(json.dumps(out), flush=True)


def cpu_info():
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return cores, model


def cpu_baseline(ctx, model, src, rec, steps_hint, n_psi):
    """FP32 CPU oracle (OpenMP over z) on the same workload, bounded to ~15-30 s: the median of 5
    timed samples on the box's CPU share (at most 16 threads, the share one GPU's job gets) is the
    value; the median of 3 one-thread samples and the host STREAM triad are reported beside it."""
    import statistics
    from oracle import pyoracle as O
    cores, cpu_model = cpu_info()
    nproc = os.cpu_count() or cores
    threads = min(cores, 16)
    O.set_threads(threads)
    o = O.Oracle3D(ctx, 32)
    o.upload_model(*model)
    o.set_receivers(rec)
    warm = o.time_steps(src, 0, 1)
    per = steps_hint if steps_hint > 0 else max(1, min(40, int(3.0 / max(warm, 1e-3))))
    samples, done = [], 1
    for _ in range(5):
        samples.append(o.time_steps(src, done, per) / per)
        done += per
    snapshot = [o.field(f) for f in range(9)]   # the state after steps 0..done-1, for the parity check
    O.set_threads(1)
    one = []
    for _ in range(3):
        one.append(o.time_steps(src, done, 1))
        done += 1
    o.close()
    O.set_threads(threads)
    triad = O.stream_triad(1 << 26, threads)   # 3 x 512 MiB of doubles: far beyond the caches
    n_upd = (ctx.gx - 8) * (ctx.gy - 8) * (ctx.gz - 8)
    step_bytes = 4.0 * (30 * n_upd + 12 * n_psi)
    t_step = statistics.median(samples)
    t_one = statistics.median(one)
    cpu_gbs = step_bytes / t_step / 1e9
    print(f"bench: cpu oracle {threads} threads: {per} steps x 5, median {t_step:.3f} s/step; 1 thread median "
          f"{t_one:.3f} s/step; host triad {triad:.1f} GB/s; nproc {nproc}", file=sys.stderr, flush=True)
    baseline = {
        "value": round(n_upd / t_step / 1e9, 5),
        "unit": "Gcell-updates/s",
        "cores": threads,
        "kind": "port",
        "host_nproc": nproc,
        "host_cpu_share": cores,
        "samples_s_per_step": [round(v, 4) for v in samples],
        "one_thread_value": round(n_upd / t_one / 1e9, 5),
        "stream_triad_gbs": round(triad, 1),
        "algorithmic_gbs": round(cpu_gbs, 2),
        "frac_of_stream_triad": round(cpu_gbs / triad, 3) if triad > 0 else None,
        "sample": f"FP32 oracle (line-faithful C++ restatement, OpenMP {threads} threads, -O3 -march=x86-64-v3, "
                  f"no FP contraction) on the same {ctx.nx}x{ctx.ny}x{ctx.nz} workload: 1 warm-up step, then 5 "
                  f"samples of {per} steps, median per step; 1 thread: median of 3 single steps; host STREAM triad "
                  f"(3 x 512 MiB doubles, {threads} threads); host: {nproc} logical CPUs, {cores} in this "
                  f"process's affinity set, {threads} used (the CPU share of one GPU's job); CPU: {cpu_model}",
    }
    return baseline, snapshot, done - 3


def snapshot_parity(ctx, model, src, oracle32, steps, tiles):
    """The bench line's own error: the 9 wavefields after `steps` steps of a fresh GPU session
    (the bench's kernels, shape `tiles`) against the FP32 oracle state of the CPU baseline and an
    FP64 oracle run, rel-L2 over the updated cells, worst field."""
    import numpy as np
    from cmakepro_cuda_seissim_v2_amd import seis3d as S
    from oracle import pyoracle as O
    g = S.GPUSimData3D(ctx)
    g.set_tiles(*tiles)
    g.upload_model(*model)
    g.run(src, 0, steps)
    gpu = [g.download_field(f) for f in range(9)]
    g.release()
    o = O.Oracle3D(ctx, 64)
    o.upload_model(*model)
    o.time_steps(src, 0, steps)
    oracle64 = [o.field(f) for f in range(9)]
    o.close()
    core = (slice(4, ctx.gz - 4), slice(4, ctx.gy - 4), slice(4, ctx.gx - 4))

    def worst(ref):
        errs = []
        for a, b in zip(gpu, ref):
            a = np.asarray(a[core], np.float64)
            b = np.asarray(b[core], np.float64)
            errs.append(float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)))
        return max(errs)

    e32, e64 = worst(oracle32), worst(oracle64)
    print(f"bench: snapshot parity after {steps} steps: rel-L2 {e32:.2e} vs FP32 oracle, {e64:.2e} vs FP64",
          file=sys.stderr, flush=True)
    return {"steps": steps, "fields": "vx vy vz sxx syy szz sxy syz sxz, updated cells",
            "snapshot_rel_l2_vs_fp32_oracle": float(f"{e32:.3e}"), "tolerance_fp32": 1e-5,
            "snapshot_rel_l2_vs_fp64_oracle": float(f"{e64:.3e}"), "tolerance_fp64": 1e-4}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note, not a measured path).", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.copy_ceiling:
        print(json.dumps(measured_ceilings()), flush=True)
        return
    if args.workload == "batch":
        if args.steps == 1000:
            args.steps = 800
        batch_bench(args)
        return
    if args.workload == "abc":
        bench_abc(args)
        return
    if args.workload == "2d":
        bench_2d(args)
        return
    if args.workload == "fwi":
        if args.steps == 1000:
            args.steps = 1200
        fwi_bench(args)
        return

    from cmakepro_cuda_seissim_v2_amd import seis3d as S
    from cmakepro_cuda_seissim_v2_amd import workloads as W

    ceilings = {} if args.no_copy_ceiling else measured_ceilings()
    nx, ny, nz = (int(v) for v in args.grid.split("x"))
    ctx = W.config2_context(nx, ny, nz)
    model = W.layered_model(ctx.nx, ctx.ny, ctx.nz)
    src = W.config2_source(ctx)
    rec = W.config2_receivers(ctx)
    run = step_run(args, S, ctx, model, src, rec, args.steps, args.warmup)
    sb, vel_us, str_us = run["bytes"], run["velocity_us"], run["stress_us"]
    vel_gbs = sb["velocity"] / (vel_us * 1e-6) / 1e9
    str_gbs = sb["stress"] / (str_us * 1e-6) / 1e9
    dominant = ("stress", str_us, sb["stress"], str_gbs) if str_us >= vel_us else ("velocity", vel_us, sb["velocity"], vel_gbs)
    family, tile_y, z_chunk = run["tiles"]
    n_upd = sb["n_upd"]
    secs = run["secs"]
    value = n_upd * args.steps / secs / 1e9
    step_gbs = sb["total"] * args.steps / secs / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(secs / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 5-layer model, Ricker force source)",
        "config": {
            "workload": f"3D O(8) velocity-stress step, {nx}x{ny}x{nz}, dx=10 m, free top + 20-cell CPML, "
                        f"15 Hz force, 150 receivers stride 2, dt=847 us (1 ms violates CFL)",
            "grid": [nx, ny, nz], "padded": [ctx.gx, ctx.gy, ctx.gz],
            "n_upd": n_upd, "n_psi": sb["n_psi"], "dt_us": round(ctx.dt * 1e6),
            "kernel_family": FAMILY_NAMES.get(family, str(family)),
            "tile_y": tile_y, "z_chunk": z_chunk, "autotuned": run["autotuned"],
            "step_hbm_gbs_algorithmic": round(step_gbs, 1),
            "step_frac_of_peak": round(step_gbs / HBM_PEAK_GBS, 4),
            "velocity_kernel_us": round(vel_us, 2), "stress_kernel_us": round(str_us, 2),
            "velocity_kernel_gbs": round(vel_gbs, 1), "stress_kernel_gbs": round(str_gbs, 1),
            "velocity_frac_of_peak": round(vel_gbs / HBM_PEAK_GBS, 4),
            "stress_frac_of_peak": round(str_gbs / HBM_PEAK_GBS, 4),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"{dominant[0]}_{FAMILY_KERNEL.get(family, 'kernel')}"
                      + ("_balanced" if family == 4 and z_chunk == 0 else ""),
            "achieved": round(dominant[3], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dominant[3] / HBM_PEAK_GBS, 4),
            "traffic": None,
            "traffic_source": None,
            "algorithmic_bytes_per_launch": dominant[2],
            "avg_launch_us": round(dominant[1], 3),
        },
        "cpu_baseline": None,
    }
    if ceilings:
        out["roofline"]["measured_ceilings_gbs"] = ceilings
        if ceilings.get("copy"):
            out["roofline"]["frac_of_copy_ceiling"] = round(dominant[3] / ceilings["copy"], 4)
    tb, src_file = profiled_traffic(out["roofline"]["kernel"], args.grid, run["tiles"])
    if tb:
        out["roofline"]["traffic"] = tb
        out["roofline"]["traffic_gbs"] = round(tb / (dominant[1] * 1e-6) / 1e9, 1)
        out["roofline"]["traffic_source"] = f"profiles/{src_file} (rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated)"
    if not args.no_cpu:
        print("bench: CPU oracle baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"], snapshot, snap_steps = cpu_baseline(ctx, model, src, rec, args.cpu_steps, sb["n_psi"])
        out["parity"] = snapshot_parity(ctx, model, src, snapshot, snap_steps, run["tiles"])
    if not args.no_extra:
        out["grid_200"] = grid_200(args, S, W)
        out["fwi"] = fwi_leg(S, W)
        out["batch"] = batch_run(800)
        out["learned_boundary"] = abc_leg()
    print(json.dumps(out), flush=True)


def step_run(args, S, ctx, model, src, rec, steps, warmup):
    """Time `steps` steps (after `warmup`), gathers recorded on the device, with the kernel shape the
    library's autotuner picks for this grid (or the one given on the command line); then per-kernel
    durations with HIP events on the session stream."""
    sb = S.step_bytes(ctx)
    g = S.GPUSimData3D(ctx)
    tuned = None
    if args.tile_y or args.z_chunk >= 0 or args.family >= 0:
        f0, ty0, zc0 = g.get_tiles()
        g.set_tiles(args.family if args.family >= 0 else f0, args.tile_y or ty0, args.z_chunk if args.z_chunk >= 0 else zc0)
    else:
        tuned = g.autotune(src, steps=20)   # times every candidate on a scratch session, keeps the best
        print(f"bench: autotuned {tuned}", file=sys.stderr, flush=True)
    tiles = g.get_tiles()
    g.upload_model(*model)
    g.set_receivers(rec)
    stride = 2
    total_steps = warmup + steps
    g.gather_begin(max(1, total_steps // stride), stride)
    print(f"bench: {ctx.nx}x{ctx.ny}x{ctx.nz} warmup {warmup} steps", file=sys.stderr, flush=True)
    g.run(src, 0, warmup)
    g.synchronize()
    t0 = time.perf_counter()
    g.run(src, warmup, steps)
    g.synchronize()
    secs = time.perf_counter() - t0
    gather = g.gather_read()
    print(f"bench: timed {steps} steps in {secs:.4f} s; gather {gather.shape}", file=sys.stderr, flush=True)
    g.gather_end()
    kt = g.run_timed(src, total_steps, args.kernel_steps)
    g.synchronize()
    g.release()
    return {"bytes": sb, "secs": secs, "tiles": tiles, "autotuned": tuned is not None,
            "velocity_us": kt.velocity_ms * 1e3 / kt.launches, "stress_us": kt.stress_ms * 1e3 / kt.launches}


def grid_200(args, S, W) -> dict:
    """The 200^3 grid with config 2's boundaries, source and receivers: throughput, per-kernel
    fraction of the HBM peak, and its error against both oracles after a bounded 41 steps."""
    import numpy as np
    from oracle import pyoracle as O
    ctx = W.config2_context(200, 200, 200)
    model = W.layered_model(200, 200, 200)
    src = W.config2_source(ctx)
    rec = W.config2_receivers(ctx)
    steps = max(args.steps, 100)
    run = step_run(args, S, ctx, model, src, rec, steps, 10)
    sb = run["bytes"]
    snap = 41
    g = S.GPUSimData3D(ctx)
    g.set_tiles(*run["tiles"])
    g.upload_model(*model)
    g.run(src, 0, snap)
    core = (slice(4, ctx.gz - 4), slice(4, ctx.gy - 4), slice(4, ctx.gx - 4))
    gpu = [g.download_field(f)[core].astype(np.float64) for f in range(9)]
    g.release()
    errs = {}
    O.set_threads(min(16, cpu_info()[0]))
    for precision in (32, 64):
        with O.Oracle3D(ctx, precision) as o:
            o.upload_model(*model)
            o.time_steps(src, 0, snap)
            errs[precision] = max(float(np.linalg.norm(a - o.field(f)[core]) / max(np.linalg.norm(o.field(f)[core]), 1e-300))
                                  for f, a in enumerate(gpu))
    out = {
        "workload": "200x200x200, dx=10 m, free top + 20-cell CPML, 15 Hz force, 150 receivers stride 2",
        "n_upd": sb["n_upd"], "n_psi": sb["n_psi"], "steps": steps,
        "gcell_updates_per_s": round(sb["n_upd"] * steps / run["secs"] / 1e9, 4),
        "ms_per_step": round(run["secs"] / steps * 1e3, 5),
        "kernel_family": FAMILY_NAMES.get(run["tiles"][0], str(run["tiles"][0])), "tile_y": run["tiles"][1],
        "z_chunk": run["tiles"][2], "autotuned": run["autotuned"],
        "velocity_kernel_us": round(run["velocity_us"], 2), "stress_kernel_us": round(run["stress_us"], 2),
        "velocity_frac_of_peak": round(sb["velocity"] / (run["velocity_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
        "stress_frac_of_peak": round(sb["stress"] / (run["stress_us"] * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
        "parity": {"steps": snap, "snapshot_rel_l2_vs_fp32_oracle": float(f"{errs[32]:.3e}"),
                   "snapshot_rel_l2_vs_fp64_oracle": float(f"{errs[64]:.3e}"),
                   "tolerance_fp32": 1e-5, "tolerance_fp64": 1e-4},
    }
    print(f"bench: grid_200 {out}", file=sys.stderr, flush=True)
    return out


def abc_leg() -> dict:
    """The learned boundaries on cases outside their training sets (64^3, free top, 360 steps, random
    source kind, position and 12-30 Hz frequency): reflected energy against a 40-cell CPML run for the
    reference's 8-cell CPML profile and the shipped learned profile (same step kernels, same cost)."""
    from cmakepro_cuda_seissim_v2_amd import learned_abc as LA
    profile = LA.load_profile()
    rows = []
    t0 = time.perf_counter()
    for seed in LA.PROFILE_HELD_OUT_SEEDS[:2]:
        case = LA.profile_case(seed)
        ref, got = LA.profile_reflection(None, case), LA.profile_reflection(profile, case)
        rows.append({"seed": seed, "frequency_hz": round(case["frequency"], 2), "cpml8_reference_profile": float(f"{ref:.4e}"),
                     "cpml8_learned_profile": float(f"{got:.4e}"), "gain": round(ref / got, 2)})
        print(f"bench: learned profile held-out {rows[-1]}", file=sys.stderr, flush=True)
    return {"workload": "8-cell CPML, learned profile vs the reference profile, 64^3 held-out cases, 360 steps, "
                        "relative squared velocity error vs 40-cell CPML, sampled every 8 steps",
            "learned_profile": list(profile), "held_out": rows, "seconds": round(time.perf_counter() - t0, 1)}


def fwi_leg(S, W) -> dict:
    """The metric's FWI half at full size: 8 shots on 200x100x200 (free top), 1200 steps,
    checkpoints every 50 steps; batched request vs the plain per-shot loop, both through
    seis3d_fwi_gradient, and the misfit."""
    import numpy as np
    steps, stride, interval = 1200, 2, 50
    ctx, start, true, shots = W.fwi_workload(8)
    observed = []
    with S.GPUSimData3D(ctx) as g:
        g.upload_model(*true)
        for src, rec in shots:
            g.reset()
            g.set_receivers(rec)
            g.gather_begin(steps // stride, stride)
            g.run(src, 0, steps)
            observed.append(g.gather_read())
            g.gather_end()
    req = [(src, rec, obs) for (src, rec), obs in zip(shots, observed)]
    with S.GPUSimData3D(ctx) as g:
        g.upload_model(*start)
        g.synchronize()
        t0 = time.perf_counter()
        batched = g.fwi_gradient(req, steps, stride, interval)
        t_batch = time.perf_counter() - t0
        t0 = time.perf_counter()
        singles = [g.fwi_gradient([r], steps, stride, interval) for r in req]
        t_plain = time.perf_counter() - t0
    summed = [np.sum([s[k].astype(np.float64) for s in singles], axis=0) for k in (1, 2, 3)]
    agree = max(float(np.linalg.norm(batched[k + 1] - summed[k]) / np.linalg.norm(summed[k])) for k in range(3))
    out = {
        "workload": "FWI gradient (vp, vs, rho), 8 shots, 200x100x200, dx=10 m, free top + 20-cell CPML, 1200 steps, "
                    "15 Hz explosions, 200 surface receivers stride 2, checkpoint every 50 steps",
        "batched_shots_per_s": round(len(shots) / t_batch, 4),
        "plain_shots_per_s": round(len(shots) / t_plain, 4),
        "speedup_vs_plain": round(t_plain / t_batch, 3),
        "batched_s": round(t_batch, 3), "plain_s": round(t_plain, 3),
        "misfit": batched[0],
        "batched_vs_summed_singles_rel_l2": float(f"{agree:.3e}"),
        "fd_checks": fwi_fd_checks(S, W, ctx, start, batched, shots, observed, steps, stride),
        "fd_tolerance": 1e-3,
    }
    print(f"bench: fwi {out}", file=sys.stderr, flush=True)
    return out


if __name__ == "__main__":
    main()
