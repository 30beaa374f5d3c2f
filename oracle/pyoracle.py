"""ctypes wrapper of the CPU oracle (oracle/liboracle3d.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check of __graft_entry__.py and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline, never as a
product path.  Contexts, sources and receivers use the boundary structs of include/seis3d.h
(the ctypes classes of cmakepro_cuda_seissim_v2_amd.seis3d), which are plain data.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from cmakepro_cuda_seissim_v2_amd.seis3d import Ctx, Receiver, Source, Survey

ORACLE_PATH = Path(__file__).resolve().parent / "lib" / "liboracle3d.so"
_LIB = None


class OracleError(RuntimeError):
    pass


class OracleInvalid(ValueError):
    pass


def library() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not ORACLE_PATH.exists():
            raise RuntimeError(f"oracle library missing: {ORACLE_PATH} (run __graft_entry__.build())")
        lib = C.CDLL(str(ORACLE_PATH))
        P = C.POINTER
        vp = C.c_void_p
        sig = {
            "oracle3d_last_error": (C.c_char_p, []),
            "oracle3d_set_threads": (C.c_int, [C.c_int]),
            "oracle3d_validate": (C.c_int, [P(Ctx)]),
            "oracle3d_coefficients": (C.c_int, [P(Ctx), C.c_int, P(C.c_float)]),
            "oracle3d_slab_count": (C.c_uint64, [P(Ctx), C.c_int]),
            "oracle3d_ricker": (C.c_double, [C.c_float, C.c_double]),
            "oracle3d_fit_survey": (C.c_int, [P(Ctx), P(Survey)]),
            "oracle3d_survey_receivers": (C.c_int, [P(Ctx), P(Survey), P(Receiver), P(C.c_int)]),
            "oracle3d_create": (C.c_int, [P(Ctx), C.c_int, P(vp)]),
            "oracle3d_destroy": (None, [vp]),
            "oracle3d_reset": (C.c_int, [vp]),
            "oracle3d_refresh_cpml": (C.c_int, [vp, P(Ctx)]),
            "oracle3d_set_profile": (C.c_int, [vp, P(C.c_double)]),
            "oracle3d_profile_coefficients": (C.c_int, [P(Ctx), P(C.c_double), C.c_int, P(C.c_float)]),
            "oracle3d_upload_plane": (C.c_int, [vp, C.c_int, P(C.c_float)]),
            "oracle3d_finish_upload": (C.c_int, [vp]),
            "oracle3d_set_receivers": (C.c_int, [vp, P(Receiver), C.c_int]),
            "oracle3d_step": (C.c_int, [vp, P(Source), C.c_uint64]),
            "oracle3d_run": (C.c_int, [vp, P(Source), C.c_uint64, C.c_int64, C.c_int, P(C.c_double), C.c_int,
                                       P(C.c_int)]),
            "oracle3d_time_steps": (C.c_double, [vp, P(Source), C.c_uint64, C.c_int64]),
            "oracle3d_sample": (C.c_int, [vp, P(C.c_double)]),
            "oracle3d_get_field": (C.c_int, [vp, C.c_int, P(C.c_double)]),
            "oracle3d_set_field": (C.c_int, [vp, C.c_int, P(C.c_double)]),
            "oracle3d_sample_scalar": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, P(C.c_double)]),
            "oracle3d_nonfinite": (C.c_int, [vp]),
            "oracle3d_stream_triad": (C.c_double, [C.c_int64, C.c_int, C.c_int]),
            "oracle3d_cpml_size": (C.c_uint64, [vp, C.c_int]),
            "oracle3d_get_cpml": (C.c_int, [vp, C.c_int, P(C.c_double)]),
            "oracle3d_set_cpml": (C.c_int, [vp, C.c_int, P(C.c_double)]),
            "oracle3d_adjoint_step": (C.c_int, [vp, P(Source), C.c_uint64, P(C.c_double)]),
            "oracle3d_fwi_gradient": (C.c_int, [vp, P(Source), C.c_int64, C.c_int, P(C.c_double), P(C.c_double),
                                                P(C.c_double), P(C.c_double), P(C.c_double)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def _check(rc: int) -> None:
    if rc == 0:
        return
    msg = library().oracle3d_last_error().decode(errors="replace")
    if rc == 1:
        raise OracleInvalid(msg)
    raise OracleError(msg)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def stream_triad(n: int, threads: int, reps: int = 5) -> float:
    """Host STREAM-triad bandwidth (GB/s) over n doubles on `threads` threads."""
    return float(library().oracle3d_stream_triad(int(n), int(threads), int(reps)))


def set_threads(n: int) -> int:
    return int(library().oracle3d_set_threads(int(n)))


def validate(ctx: Ctx) -> Ctx:
    _check(library().oracle3d_validate(C.byref(ctx)))
    return ctx


def coefficients(ctx: Ctx, axis: int) -> np.ndarray:
    n = (ctx.gx, ctx.gy, ctx.gz)[axis]
    out = np.zeros(n * 6, np.float32)
    _check(library().oracle3d_coefficients(C.byref(ctx), axis, out.ctypes.data_as(C.POINTER(C.c_float))))
    return out


def profile_coefficients(ctx: Ctx, profile, axis: int) -> np.ndarray:
    """Learned-profile CPML table of one axis, [half][n][a, b, 1/kappa] (FP32)."""
    n = (ctx.gx, ctx.gy, ctx.gz)[axis]
    out = np.zeros(n * 6, np.float32)
    p = np.ascontiguousarray(profile, dtype=np.float64)
    _check(library().oracle3d_profile_coefficients(C.byref(ctx), _dptr(p), axis,
                                                   out.ctypes.data_as(C.POINTER(C.c_float))))
    return out


def slab_count(ctx: Ctx, axis: int) -> int:
    return int(library().oracle3d_slab_count(C.byref(ctx), axis))


def ricker(frequency: float, time: float) -> float:
    return float(library().oracle3d_ricker(frequency, time))


def fit_survey(ctx: Ctx, survey: Survey) -> Survey:
    _check(library().oracle3d_fit_survey(C.byref(ctx), C.byref(survey)))
    return survey


def survey_receivers(ctx: Ctx, survey: Survey) -> list[tuple[int, int, int]]:
    buf = (Receiver * 300)()
    n = C.c_int(300)
    _check(library().oracle3d_survey_receivers(C.byref(ctx), C.byref(survey), buf, C.byref(n)))
    return [(buf[i].x, buf[i].y, buf[i].z) for i in range(n.value)]


class Oracle3D:
    """CPU restatement of one Seis3D session, FP32 (precision=32) or FP64 (precision=64)."""

    def __init__(self, ctx: Ctx, precision: int = 32):
        self.ctx = ctx
        self.precision = precision
        self._h = C.c_void_p()
        _check(library().oracle3d_create(C.byref(ctx), precision, C.byref(self._h)))

    def close(self) -> None:
        if self._h.value:
            library().oracle3d_destroy(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def reset(self) -> None:
        _check(library().oracle3d_reset(self._h))

    def refresh_cpml(self, ctx: Ctx) -> None:
        _check(library().oracle3d_refresh_cpml(self._h, C.byref(ctx)))

    def set_profile(self, profile) -> None:
        """Learned CPML profile tables (None: the reference profile)."""
        if profile is None:
            _check(library().oracle3d_set_profile(self._h, None))
            return
        p = np.ascontiguousarray(profile, dtype=np.float64)
        _check(library().oracle3d_set_profile(self._h, _dptr(p)))

    def upload_plane(self, z: int, vp_vs_rho: np.ndarray) -> None:
        a = np.ascontiguousarray(vp_vs_rho, dtype=np.float32)
        _check(library().oracle3d_upload_plane(self._h, int(z), a.ctypes.data_as(C.POINTER(C.c_float))))

    def finish_upload(self) -> None:
        _check(library().oracle3d_finish_upload(self._h))

    def upload_model(self, vp: np.ndarray, vs: np.ndarray, rho: np.ndarray) -> None:
        for z in range(self.ctx.nz):
            self.upload_plane(z, np.stack([vp[z], vs[z], rho[z]], axis=-1).astype(np.float32))
        self.finish_upload()

    def set_receivers(self, receivers) -> None:
        rec = list(receivers)
        arr = (Receiver * max(1, len(rec)))()
        for i, r in enumerate(rec):
            arr[i] = Receiver(int(r[0]), int(r[1]), int(r[2]))
        _check(library().oracle3d_set_receivers(self._h, arr, len(rec)))
        self.n_receivers = len(rec)

    def step(self, src: Source, step: int) -> None:
        _check(library().oracle3d_step(self._h, C.byref(src), int(step)))

    def run(self, src: Source, first_step: int, n_steps: int, stride: int = 0, max_samples: int = 0):
        """Run n_steps; returns the [sample][receiver][3] gather when stride > 0."""
        nrec = getattr(self, "n_receivers", 0)
        if stride > 0 and nrec:
            out = np.zeros(max(1, max_samples) * nrec * 3)
            got = C.c_int()
            _check(library().oracle3d_run(self._h, C.byref(src), int(first_step), int(n_steps), int(stride),
                                          _dptr(out), int(max_samples), C.byref(got)))
            return out[: got.value * nrec * 3].reshape(got.value, nrec, 3)
        _check(library().oracle3d_run(self._h, C.byref(src), int(first_step), int(n_steps), 0, None, 0, None))
        return None

    def time_steps(self, src: Source, first_step: int, n_steps: int) -> float:
        return float(library().oracle3d_time_steps(self._h, C.byref(src), int(first_step), int(n_steps)))

    def sample(self) -> np.ndarray:
        n = getattr(self, "n_receivers", 0)
        out = np.zeros(max(1, n * 3))
        _check(library().oracle3d_sample(self._h, _dptr(out)))
        return out[: n * 3].reshape(n, 3)

    def field(self, f: int) -> np.ndarray:
        c = self.ctx
        out = np.zeros(c.gz * c.gy * c.gx)
        _check(library().oracle3d_get_field(self._h, int(f), _dptr(out)))
        return out.reshape(c.gz, c.gy, c.gx)

    def set_field(self, f: int, values: np.ndarray) -> None:
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        _check(library().oracle3d_set_field(self._h, int(f), _dptr(a)))

    def cpml(self, axis: int) -> np.ndarray:
        """CPML memories of one axis (all slots, slab layout)."""
        out = np.zeros(int(library().oracle3d_cpml_size(self._h, int(axis))))
        _check(library().oracle3d_get_cpml(self._h, int(axis), _dptr(out)))
        return out

    def set_cpml(self, axis: int, values: np.ndarray) -> None:
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        if a.size != int(library().oracle3d_cpml_size(self._h, int(axis))):
            raise ValueError("CPML memory size mismatch")
        _check(library().oracle3d_set_cpml(self._h, int(axis), _dptr(a)))

    def adjoint_step(self, src: Source, step: int, fields, cpml):
        """Adjoint of step `step` applied to an adjoint state (9 padded wavefields, 3 CPML arrays).

        Returns the new (fields, cpml); the model-gradient sums are discarded.
        """
        sizes = [int(library().oracle3d_cpml_size(self._h, a)) for a in range(3)]
        flat = np.concatenate([np.asarray(f, np.float64).reshape(-1) for f in fields] +
                              [np.asarray(c, np.float64).reshape(-1) for c in cpml])
        cells = self.ctx.gx * self.ctx.gy * self.ctx.gz
        if flat.size != 9 * cells + sum(sizes):
            raise ValueError("adjoint state size mismatch")
        _check(library().oracle3d_adjoint_step(self._h, C.byref(src), int(step), _dptr(flat)))
        shape = (self.ctx.gz, self.ctx.gy, self.ctx.gx)
        out_f = [flat[k * cells:(k + 1) * cells].reshape(shape).copy() for k in range(9)]
        o, out_c = 9 * cells, []
        for n in sizes:
            out_c.append(flat[o:o + n].copy())
            o += n
        return out_f, out_c

    def sample_scalar(self, x: int, y: int, z: int, component: int) -> float:
        out = C.c_double()
        _check(library().oracle3d_sample_scalar(self._h, x, y, z, component, C.byref(out)))
        return out.value

    def fwi_gradient(self, src: Source, n_steps: int, stride: int, observed: np.ndarray):
        """Misfit and its gradient w.r.t. vp, vs, rho ([nz, ny, nx]) by the discrete adjoint."""
        c = self.ctx
        obs = np.ascontiguousarray(observed, dtype=np.float64).reshape(-1)
        shape = (c.nz, c.ny, c.nx)
        g = [np.zeros(shape) for _ in range(3)]
        m = C.c_double()
        _check(library().oracle3d_fwi_gradient(self._h, C.byref(src), int(n_steps), int(stride), _dptr(obs),
                                               C.byref(m), *[_dptr(a) for a in g]))
        return m.value, g[0], g[1], g[2]

    @property
    def nonfinite(self) -> bool:
        return bool(library().oracle3d_nonfinite(self._h))
