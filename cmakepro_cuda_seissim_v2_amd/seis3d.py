"""Python host mirror of the Seis3D solver-session API over the C ABI (include/seis3d.h).

Names follow the reference's operator interface (cuda/Cuda_Seis3D.cuh:6-42):
``validateContext3D``, ``estimateMemory3D``, ``initializeGpu3D``, ``releaseGpu3D``,
``resetGpu3D``, ``refreshCpmlCoefficients3D``, ``runStep3D``, ``setReceivers3D``,
``sampleReceivers3D``, ``probeMaterial3D``, ``uploadModelPlane3D``,
``finishModelUpload3D``, ``downloadModelPlane3D``, and the shot-record and model files
``exportGather3D``, ``readShotSegy3D``, ``inspectModel3D``, ``importModel3D``,
``exportModel3D`` (src/Seis3D/Acquisition3D.cpp, ShotRecordIo3D.cpp, ModelIo3D.cpp); error behaviour maps the C codes
back to the reference's exception classes (invalid_argument -> ``InvalidArgument``
(a ValueError), runtime_error -> ``SolverRuntimeError``, logic_error -> ``LogicError``).

All compute runs in the HIP library ``lib/libseis3d.so``; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libseis3d.so"

VX, VY, VZ, SXX, SYY, SZZ, SXY, SYZ, SXZ = range(9)
INV_RHO, LAMBDA, MU = 9, 10, 11
EXPLOSION, FORCE = 0, 1
(UNIFORM, TWO_LAYER, SPHERE, IMPORTED, THREE_LAYER, DIPPING_LAYER, LOW_VELOCITY_LAYER, FAULT,
 GRADIENT, SALT_DOME) = range(10)
MAX_RECEIVERS = 300


class InvalidArgument(ValueError):
    """std::invalid_argument of the reference (error code 1)."""


class SolverRuntimeError(RuntimeError):
    """std::runtime_error of the reference (error codes 2 and 4)."""


class LogicError(RuntimeError):
    """std::logic_error of the reference (error code 3)."""


class NonFiniteError(SolverRuntimeError):
    """Non-finite wavefield detected (error code 4)."""


class Ctx(C.Structure):
    """seis3d_ctx = SimulationContext3D (include/Common3D.h:24-42)."""
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("pml", C.c_int32),
        ("free_surface", C.c_int32), ("reflecting", C.c_int32),
        ("spacing", C.c_float), ("cfl", C.c_float), ("dt", C.c_float),
        ("manual_dt", C.c_int32), ("requested_dt_ms", C.c_float), ("safe_dt", C.c_float),
        ("vp", C.c_float), ("vs", C.c_float), ("density", C.c_float),
        ("layer_vp", C.c_float), ("layer_vs", C.c_float), ("layer_density", C.c_float),
        ("max_vp", C.c_float), ("frequency", C.c_float),
        ("model_kind", C.c_int32),
        ("interface_depth", C.c_float), ("layer_dip", C.c_float), ("layer_thickness", C.c_float),
        ("fault_throw", C.c_float), ("anomaly_radius", C.c_float),
        ("gx", C.c_int32), ("gy", C.c_int32), ("gz", C.c_int32),
        ("ox", C.c_int32), ("oy", C.c_int32), ("oz", C.c_int32),
        ("cells", C.c_uint64),
    ]

    def copy(self) -> "Ctx":
        out = Ctx()
        C.pointer(out)[0] = self
        return out

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Source(C.Structure):
    """seis3d_source = Source3D (include/Common3D.h:44-49)."""
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32), ("kind", C.c_int32),
                ("amplitude", C.c_float), ("direction", C.c_float * 3)]


class Receiver(C.Structure):
    """seis3d_receiver = Receiver3D (include/Common3D.h:51)."""
    _fields_ = [("x", C.c_int32), ("y", C.c_int32), ("z", C.c_int32)]


class Budget(C.Structure):
    """seis3d_budget = MemoryBudget3D (include/Common3D.h:72-74)."""
    _fields_ = [("field_bytes", C.c_uint64), ("cpml_bytes", C.c_uint64), ("view_bytes", C.c_uint64),
                ("total_bytes", C.c_uint64)]


class Survey(C.Structure):
    """seis3d_survey = the spread part of Survey3D (include/Seis3D_Sim.h:36-45)."""
    _fields_ = [("receiver_count_x", C.c_int32), ("receiver_count_y", C.c_int32),
                ("receiver_origin", C.c_int32 * 3), ("receiver_spacing", C.c_int32 * 2),
                ("shot_count_x", C.c_int32), ("shot_count_y", C.c_int32),
                ("shot_origin", C.c_int32 * 3), ("shot_spacing", C.c_int32 * 2)]


class Shot(C.Structure):
    """seis3d_shot: source, receivers and observed [samples][receivers][3] data of one FWI shot."""
    _fields_ = [("source", Source), ("receivers", C.POINTER(Receiver)), ("receiver_count", C.c_int32),
                ("observed", C.POINTER(C.c_float))]


class GatherInfo(C.Structure):
    """seis3d_gather_info: frozen acquisition metadata of one shot gather (Gather3D)."""
    _fields_ = [("spacing", C.c_float), ("frequency", C.c_float), ("source", Source),
                ("world_origin", C.c_double * 2), ("cmp_bin", C.c_float * 2), ("shot_id", C.c_int32),
                ("stride", C.c_int32), ("samples", C.c_int32), ("sample_interval_us", C.c_int32),
                ("coordinate_scalar", C.c_int32), ("first_sample_time", C.c_double)]


class SegyInfo(C.Structure):
    """seis3d_segy_info: header scan of one single-component shot file."""
    _fields_ = [("traces", C.c_int32), ("samples", C.c_int32), ("sample_interval_us", C.c_int32),
                ("format", C.c_int32), ("revision", C.c_int32), ("component", C.c_int32),
                ("shot_id", C.c_int32), ("measurement_unit", C.c_int32), ("coordinate_units", C.c_int32)]


class ViewLayout(C.Structure):
    """seis3d_view_layout: ViewLayout3D sample counts of the packed view."""
    _fields_ = [("slice_offset", C.c_uint64 * 3), ("slice_width", C.c_int32 * 3), ("slice_height", C.c_int32 * 3),
                ("volume_offset", C.c_uint64), ("total_samples", C.c_uint64), ("volume_x", C.c_int32),
                ("volume_y", C.c_int32), ("volume_z", C.c_int32), ("domain_size", C.c_int32 * 3),
                ("domain_origin", C.c_int32 * 3)]


class ViewStats(C.Structure):
    _fields_ = [("maximum", C.c_float), ("material_min", C.c_float), ("material_max", C.c_float),
                ("velocity_maximum", C.c_float)]


class Timing(C.Structure):
    _fields_ = [("velocity_ms", C.c_double), ("stress_ms", C.c_double), ("total_ms", C.c_double),
                ("launches", C.c_int64)]


_LIB = None


def library() -> C.CDLL:
    """Load the HIP library; raises if it has not been built (no fallback)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"HIP solver library missing: {LIB_PATH} (run __graft_entry__.build())")
        lib = C.CDLL(str(LIB_PATH))
        P = C.POINTER
        vp = C.c_void_p
        sig = {
            "seis3d_default_ctx": (None, [P(Ctx)]),
            "seis3d_default_source": (None, [P(Source)]),
            "seis3d_default_survey": (None, [P(Survey)]),
            "seis3d_validate_ctx": (C.c_int, [P(Ctx)]),
            "seis3d_estimate_memory": (C.c_int, [P(Ctx), P(Budget)]),
            "seis3d_device_bytes": (C.c_int, [P(Ctx), P(C.c_uint64)]),
            "seis3d_cpml_coefficients": (C.c_int, [P(Ctx), C.c_int, P(C.c_float), C.c_size_t]),
            "seis3d_fit_survey": (C.c_int, [P(Ctx), P(Survey)]),
            "seis3d_survey_receivers": (C.c_int, [P(Ctx), P(Survey), P(Receiver), P(C.c_int)]),
            "seis3d_sequence_shot": (C.c_int, [P(Ctx), P(Survey), C.c_int, P(Source)]),
            "seis3d_gather_timing": (C.c_int, [P(Ctx), C.c_int, P(C.c_int), P(C.c_double)]),
            "seis3d_ricker": (C.c_float, [C.c_float, C.c_double]),
            "seis3d_create": (C.c_int, [P(Ctx), P(vp)]),
            "seis3d_batch_create_mode": (C.c_int, [P(vp), P(Source), C.c_int, C.c_int64, C.c_int, P(vp)]),
            "seis3d_bandwidth_probe": (C.c_int, [C.c_int, C.c_uint64, C.c_int, P(C.c_double), P(C.c_double)]),
            "seis3d_destroy": (None, [vp]),
            "seis3d_get_ctx": (C.c_int, [vp, P(Ctx)]),
            "seis3d_reset": (C.c_int, [vp]),
            "seis3d_refresh_cpml": (C.c_int, [vp, P(Ctx)]),
            "seis3d_read_cpml": (C.c_int, [vp, C.c_int, P(C.c_float), C.c_size_t]),
            "seis3d_set_cpml_profile": (C.c_int, [vp, P(C.c_double)]),
            "seis3d_get_cpml_profile": (C.c_int, [vp, P(C.c_double), P(C.c_int)]),
            "seis3d_cpml_profile_coefficients": (C.c_int, [P(Ctx), P(C.c_double), C.c_int, P(C.c_float),
                                                           C.c_size_t]),
            "seis3d_upload_plane": (C.c_int, [vp, C.c_int, P(C.c_float)]),
            "seis3d_finish_upload": (C.c_int, [vp]),
            "seis3d_download_plane": (C.c_int, [vp, C.c_int, P(C.c_float)]),
            "seis3d_probe": (C.c_int, [vp, P(Receiver), P(C.c_float)]),
            "seis3d_step": (C.c_int, [vp, P(Source), C.c_uint64]),
            "seis3d_run": (C.c_int, [vp, P(Source), C.c_uint64, C.c_int64]),
            "seis3d_synchronize": (C.c_int, [vp]),
            "seis3d_set_receivers": (C.c_int, [vp, P(Receiver), C.c_int]),
            "seis3d_receiver_count": (C.c_int, [vp]),
            "seis3d_sample_receivers": (C.c_int, [vp, P(P(C.c_float))]),
            "seis3d_gather_begin": (C.c_int, [vp, C.c_int, C.c_int]),
            "seis3d_gather_count": (C.c_int, [vp]),
            "seis3d_gather_read": (C.c_int, [vp, P(C.c_float), C.c_size_t, P(C.c_int)]),
            "seis3d_gather_end": (C.c_int, [vp]),
            "seis3d_download_field": (C.c_int, [vp, C.c_int, P(C.c_float)]),
            "seis3d_upload_field": (C.c_int, [vp, C.c_int, P(C.c_float)]),
            "seis3d_extract_slice": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, P(C.c_float), C.c_size_t]),
            "seis3d_set_tiles": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
            "seis3d_make_view_layout": (C.c_int, [P(Ctx), P(ViewLayout)]),
            "seis3d_extract_views": (C.c_int, [vp, P(C.c_float), C.c_size_t, P(C.c_int32), C.c_int, C.c_int, C.c_int,
                                               C.c_int, P(ViewStats)]),
            "seis3d_extract_views_noalloc": (C.c_int, [vp, P(C.c_float), C.c_size_t, P(C.c_int32), C.c_int, C.c_int,
                                                       C.c_int, C.c_int, P(ViewStats)]),
            "seis3d_get_tiles": (C.c_int, [vp, P(C.c_int), P(C.c_int), P(C.c_int)]),
            "seis3d_run_timed": (C.c_int, [vp, P(Source), C.c_uint64, C.c_int64, P(Timing)]),
            "seis3d_step_bytes": (C.c_int, [P(Ctx), P(C.c_uint64), P(C.c_uint64), P(C.c_uint64),
                                            P(C.c_uint64), P(C.c_uint64)]),
            "seis3d_last_error": (C.c_char_p, []),
            "seis3d_version": (C.c_char_p, []),
            "seis3d_device_count": (C.c_int, []),
            "seis3d_autotune": (C.c_int, [vp, P(Source), C.c_int, P(C.c_int), P(C.c_int), P(C.c_int),
                                          P(C.c_double)]),
            "seis3d_batch_create": (C.c_int, [P(vp), P(Source), C.c_int, C.c_int64, P(vp)]),
            "seis3d_batch_launch": (C.c_int, [vp]),
            "seis3d_batch_synchronize": (C.c_int, [vp]),
            "seis3d_batch_destroy": (None, [vp]),
            "seis3d_abc_param_count": (C.c_int, []),
            "seis3d_abc_attach": (C.c_int, [vp, P(C.c_float)]),
            "seis3d_abc_detach": (C.c_int, [vp]),
            "seis3d_abc_frame_cells": (C.c_int, [vp, P(C.c_int)]),
            "seis3d_copy_wavefield_box": (C.c_int, [vp, vp, C.c_int, C.c_int, C.c_int]),
            "seis3d_abc_trainer_create": (C.c_int, [C.c_int, C.c_uint64, P(vp)]),
            "seis3d_abc_trainer_destroy": (None, [vp]),
            "seis3d_abc_trainer_count": (C.c_int, [vp]),
            "seis3d_abc_trainer_add": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                                 P(C.c_int)]),
            "seis3d_abc_train": (C.c_int, [vp, P(C.c_float), C.c_int, C.c_int, C.c_float, P(C.c_float)]),
            "seis3d_abc_evaluate": (C.c_int, [vp, P(C.c_float), P(C.c_double)]),
            "seis3d_abc_mlp": (C.c_int, [P(C.c_float), P(C.c_float), P(C.c_float), C.c_int, P(C.c_float),
                                         P(C.c_float), P(C.c_float), P(C.c_float), P(C.c_double)]),
            "seis3d_gather_info_init": (C.c_int, [P(Ctx), P(Source), P(Receiver), C.c_int, C.c_int, C.c_int,
                                                  C.c_int, P(C.c_double), P(C.c_float), P(GatherInfo)]),
            "seis3d_export_gather": (C.c_int, [P(GatherInfo), P(Receiver), C.c_int, P(C.c_float), C.c_char_p,
                                               C.c_char_p]),
            "seis3d_segy_inspect": (C.c_int, [C.c_char_p, P(SegyInfo)]),
            "seis3d_segy_read": (C.c_int, [C.c_char_p, P(C.c_float), P(C.c_double), P(C.c_double),
                                           P(C.c_int32)]),
            "seis3d_s3dm_inspect": (C.c_int, [C.c_char_p, P(Ctx)]),
            "seis3d_s3dm_load": (C.c_int, [vp, C.c_char_p]),
            "seis3d_s3dm_save": (C.c_int, [vp, C.c_char_p]),
            "seis3d_s3dm_write": (C.c_int, [C.c_char_p, P(Ctx), P(C.c_float)]),
            "seis3d_fwi_plan": (C.c_int, [P(Ctx), C.c_int64, C.c_int, C.c_int, C.c_int, P(C.c_uint64),
                                          P(C.c_int)]),
            "seis3d_fwi_gradient": (C.c_int, [vp, P(Shot), C.c_int, C.c_int64, C.c_int, C.c_int, P(C.c_double),
                                              P(C.c_float), P(C.c_float), P(C.c_float)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


def exported_symbols() -> list[str]:
    """C symbols this wrapper binds (must match include/seis3d.h)."""
    library()
    return sorted(n for n in dir(_LIB) if n.startswith("seis3d_"))


def _check(rc: int) -> None:
    if rc == 0:
        return
    msg = library().seis3d_last_error().decode(errors="replace")
    if rc == 1:
        raise InvalidArgument(msg)
    if rc == 3:
        raise LogicError(msg)
    if rc == 4:
        raise NonFiniteError(msg)
    raise SolverRuntimeError(msg)


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


# ---- context helpers -------------------------------------------------------

def default_context(**overrides) -> Ctx:
    """SimulationContext3D{} defaults (include/Common3D.h:24-42) with overrides."""
    c = Ctx()
    library().seis3d_default_ctx(C.byref(c))
    for k, v in overrides.items():
        if not hasattr(c, k):
            raise AttributeError(k)
        setattr(c, k, v)
    return c


def default_source(**overrides) -> Source:
    s = Source()
    library().seis3d_default_source(C.byref(s))
    for k, v in overrides.items():
        if k == "direction":
            for i in range(3):
                s.direction[i] = float(v[i])
        else:
            setattr(s, k, v)
    return s


def default_survey(**overrides) -> Survey:
    q = Survey()
    library().seis3d_default_survey(C.byref(q))
    for k, v in overrides.items():
        f = getattr(q, k)
        if isinstance(f, C.Array):
            for i, x in enumerate(v):
                f[i] = x
        else:
            setattr(q, k, v)
    return q


def validateContext3D(ctx: Ctx) -> Ctx:
    """validateContext3D (cuda/Cuda_Seis3D.cu:433-483): derives dims and dt in place."""
    _check(library().seis3d_validate_ctx(C.byref(ctx)))
    return ctx


def estimateMemory3D(ctx: Ctx) -> Budget:
    b = Budget()
    _check(library().seis3d_estimate_memory(C.byref(ctx), C.byref(b)))
    return b


def device_bytes(ctx: Ctx) -> int:
    v = C.c_uint64()
    _check(library().seis3d_device_bytes(C.byref(ctx), C.byref(v)))
    return int(v.value)


def makeCoefficients3D(ctx: Ctx, axis: int) -> np.ndarray:
    n = (ctx.gx, ctx.gy, ctx.gz)[axis]
    out = np.zeros(n * 6, np.float32)
    _check(library().seis3d_cpml_coefficients(C.byref(ctx), axis, _fptr(out), out.size))
    return out


PROFILE_PARAMS = 5
REFERENCE_PROFILE = (0.0, 3.0, 3.0, 1.0, 1.0)   # makeCoefficients3D's own shapes


def _profile_ptr(profile):
    if profile is None:
        return None, None
    p = np.ascontiguousarray(profile, dtype=np.float64)
    if p.shape != (PROFILE_PARAMS,):
        raise InvalidArgument("Seis3D: a CPML profile has 5 parameters.")
    return p, p.ctypes.data_as(C.POINTER(C.c_double))


def cpml_profile_coefficients(ctx: Ctx, profile, axis: int) -> np.ndarray:
    """[half][n][a, b, 1/kappa] of one axis for a learned CPML profile (None: the reference's)."""
    n = (ctx.gx, ctx.gy, ctx.gz)[axis]
    out = np.zeros(n * 6, np.float32)
    keep, ptr = _profile_ptr(profile)
    _check(library().seis3d_cpml_profile_coefficients(C.byref(ctx), ptr, axis, _fptr(out), out.size))
    return out


def fitSurvey3D(ctx: Ctx, survey: Survey | None = None) -> Survey:
    q = survey if survey is not None else default_survey()
    _check(library().seis3d_fit_survey(C.byref(ctx), C.byref(q)))
    return q


def configureReceivers3D(ctx: Ctx, survey: Survey) -> list[tuple[int, int, int]]:
    buf = (Receiver * MAX_RECEIVERS)()
    n = C.c_int(MAX_RECEIVERS)
    _check(library().seis3d_survey_receivers(C.byref(ctx), C.byref(survey), buf, C.byref(n)))
    return [(buf[i].x, buf[i].y, buf[i].z) for i in range(n.value)]


def positionSequenceShot3D(ctx: Ctx, survey: Survey, index: int, src: Source) -> Source:
    _check(library().seis3d_sequence_shot(C.byref(ctx), C.byref(survey), index, C.byref(src)))
    return src


def gather_timing(ctx: Ctx, stride: int) -> tuple[int, float]:
    iv = C.c_int()
    t0 = C.c_double()
    _check(library().seis3d_gather_timing(C.byref(ctx), stride, C.byref(iv), C.byref(t0)))
    return iv.value, t0.value


def step_bytes(ctx: Ctx) -> dict:
    vals = [C.c_uint64() for _ in range(5)]
    _check(library().seis3d_step_bytes(C.byref(ctx), *[C.byref(v) for v in vals]))
    keys = ("total", "velocity", "stress", "n_upd", "n_psi")
    return {k: int(v.value) for k, v in zip(keys, vals)}


def ricker(frequency: float, time: float) -> float:
    return float(library().seis3d_ricker(frequency, time))


PROBE_KINDS = {"read": 0, "write": 1, "copy": 2, "copy_lds_dma": 3, "mix10_3": 4, "mix12_6": 5}


def bandwidth_probe(kind: str, mib: int = 1024, reps: int = 10) -> dict:
    """Project-written HBM stream kernels (the measured ceiling): GB/s and microseconds per launch."""
    gbs, us = C.c_double(), C.c_double()
    _check(library().seis3d_bandwidth_probe(PROBE_KINDS[kind], int(mib) << 20, int(reps), C.byref(gbs), C.byref(us)))
    return {"kind": kind, "mib_per_array": mib, "gbs": gbs.value, "us_per_launch": us.value}


def version() -> str:
    return library().seis3d_version().decode()


# ---- session ---------------------------------------------------------------

class GPUSimData3D:
    """One solver session (GPUSimData3D of include/Common3D.h:76-93) on the current HIP device."""

    def __init__(self, ctx: Ctx | None = None):
        self._h = C.c_void_p()
        self.ctx: Ctx | None = None
        if ctx is not None:
            self.initialize(ctx)

    # initializeGpu3D (cuda/Cuda_Seis3D.cu:519-555)
    def initialize(self, ctx: Ctx) -> "GPUSimData3D":
        if self._h.value:
            raise LogicError("Seis3D: release the previous session before initializing.")
        h = C.c_void_p()
        _check(library().seis3d_create(C.byref(ctx), C.byref(h)))
        self._h = h
        self.ctx = ctx.copy()
        return self

    # releaseGpu3D (cuda/Cuda_Seis3D.cu:558-569): idempotent
    def release(self) -> None:
        if self._h.value:
            library().seis3d_destroy(self._h)
        self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.release()

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def _ctx_now(self) -> Ctx:
        c = Ctx()
        _check(library().seis3d_get_ctx(self._h, C.byref(c)))
        return c

    def reset(self) -> None:
        _check(library().seis3d_reset(self._h))

    def refresh_cpml(self, ctx: Ctx) -> None:
        _check(library().seis3d_refresh_cpml(self._h, C.byref(ctx)))
        self.ctx = self._ctx_now()

    def set_cpml_profile(self, profile) -> None:
        """Use a learned CPML profile (5 parameters) for this session's tables; None restores the
        reference profile."""
        keep, ptr = _profile_ptr(profile)
        _check(library().seis3d_set_cpml_profile(self._h, ptr))

    def cpml_profile(self):
        """(profile parameters, learned?) of this session's CPML tables."""
        out = np.zeros(PROFILE_PARAMS, np.float64)
        learned = C.c_int()
        _check(library().seis3d_get_cpml_profile(self._h, out.ctypes.data_as(C.POINTER(C.c_double)),
                                                 C.byref(learned)))
        return out, bool(learned.value)

    def read_cpml(self, axis: int) -> np.ndarray:
        n = (self.ctx.gx, self.ctx.gy, self.ctx.gz)[axis]
        out = np.zeros(n * 6, np.float32)
        _check(library().seis3d_read_cpml(self._h, axis, _fptr(out), out.size))
        return out

    def upload_plane(self, z: int, vp_vs_rho: np.ndarray) -> None:
        a = np.ascontiguousarray(vp_vs_rho, dtype=np.float32)
        if a.size != self.ctx.nx * self.ctx.ny * 3:
            raise InvalidArgument("Seis3D: invalid model plane.")
        _check(library().seis3d_upload_plane(self._h, int(z), _fptr(a)))

    def finish_upload(self) -> None:
        _check(library().seis3d_finish_upload(self._h))

    def upload_model(self, vp: np.ndarray, vs: np.ndarray, rho: np.ndarray) -> None:
        """Upload a whole physical model [nz, ny, nx] plane by plane, then extend."""
        for z in range(self.ctx.nz):
            plane = np.stack([vp[z], vs[z], rho[z]], axis=-1).astype(np.float32)
            self.upload_plane(z, plane)
        self.finish_upload()

    def download_plane(self, z: int) -> np.ndarray:
        out = np.zeros(self.ctx.nx * self.ctx.ny * 3, np.float32)
        _check(library().seis3d_download_plane(self._h, int(z), _fptr(out)))
        return out.reshape(self.ctx.ny, self.ctx.nx, 3)

    def probe(self, x: int, y: int, z: int) -> np.ndarray:
        vals = np.zeros(3, np.float32)
        p = Receiver(x, y, z)
        _check(library().seis3d_probe(self._h, C.byref(p), _fptr(vals)))
        return vals

    def step(self, src: Source, step: int) -> None:
        _check(library().seis3d_step(self._h, C.byref(src), int(step)))

    def run(self, src: Source, first_step: int, n_steps: int) -> None:
        _check(library().seis3d_run(self._h, C.byref(src), int(first_step), int(n_steps)))

    def run_timed(self, src: Source, first_step: int, n_steps: int) -> Timing:
        t = Timing()
        _check(library().seis3d_run_timed(self._h, C.byref(src), int(first_step), int(n_steps), C.byref(t)))
        return t

    def synchronize(self) -> None:
        _check(library().seis3d_synchronize(self._h))

    def set_receivers(self, receivers) -> None:
        rec = list(receivers)
        arr = (Receiver * max(1, len(rec)))()
        for i, r in enumerate(rec):
            arr[i] = Receiver(int(r[0]), int(r[1]), int(r[2]))
        _check(library().seis3d_set_receivers(self._h, arr if rec else None, len(rec)))

    @property
    def receiver_count(self) -> int:
        return int(library().seis3d_receiver_count(self._h))

    def sample_receivers(self) -> np.ndarray | None:
        ptr = C.POINTER(C.c_float)()
        _check(library().seis3d_sample_receivers(self._h, C.byref(ptr)))
        if not ptr:
            return None
        n = self.receiver_count
        return np.ctypeslib.as_array(ptr, shape=(n * 3,)).copy().reshape(n, 3)

    def gather_begin(self, target_samples: int, stride: int) -> None:
        _check(library().seis3d_gather_begin(self._h, int(target_samples), int(stride)))

    @property
    def gather_count(self) -> int:
        return int(library().seis3d_gather_count(self._h))

    def gather_read(self) -> np.ndarray:
        n = self.gather_count
        r = self.receiver_count
        out = np.zeros(max(1, n * r * 3), np.float32)
        got = C.c_int()
        _check(library().seis3d_gather_read(self._h, _fptr(out), out.size, C.byref(got)))
        return out[: got.value * r * 3].reshape(got.value, r, 3)

    def gather_end(self) -> None:
        _check(library().seis3d_gather_end(self._h))

    def download_field(self, field: int) -> np.ndarray:
        c = self.ctx
        out = np.zeros(c.gz * c.gy * c.gx, np.float32)
        _check(library().seis3d_download_field(self._h, int(field), _fptr(out)))
        return out.reshape(c.gz, c.gy, c.gx)

    def upload_field(self, field: int, values: np.ndarray) -> None:
        a = np.ascontiguousarray(values, dtype=np.float32).reshape(-1)
        c = self.ctx
        if a.size != c.gx * c.gy * c.gz:
            raise InvalidArgument("Seis3D: invalid field size.")
        _check(library().seis3d_upload_field(self._h, int(field), _fptr(a)))

    def extract_slice(self, axis: int, at: int, component: int) -> np.ndarray:
        c = self.ctx
        pml = 0 if c.reflecting else c.pml
        dom = (c.nx + 2 * pml, c.ny + 2 * pml, c.nz + (1 if c.free_surface else 2) * pml)
        u = 1 if axis == 0 else 0
        v = 1 if axis == 2 else 2
        out = np.zeros(dom[u] * dom[v], np.float32)
        _check(library().seis3d_extract_slice(self._h, axis, at, component, _fptr(out), out.size))
        return out.reshape(dom[v], dom[u])

    def extract_views(self, slices, component: int = 0, material_component: int = 5, volume: bool = True,
                      include_boundary: bool = False, no_alloc: bool = False):
        """extractViews3D: ([sample][wave, material] array, statistics dict).  no_alloc: through
        seis3d_extract_views_noalloc (the library's fixed per-GPU staging; no device allocation)."""
        lay = make_view_layout(self.ctx)
        n = lay.total_samples if volume else lay.volume_offset
        out = np.zeros(2 * n, np.float32)
        sl = (C.c_int32 * 3)(*[int(v) for v in slices])
        st = ViewStats()
        fn = library().seis3d_extract_views_noalloc if no_alloc else library().seis3d_extract_views
        _check(fn(self._h, _fptr(out), out.size, sl, int(component), int(material_component), int(volume),
                  int(include_boundary), C.byref(st)))
        return out.reshape(n, 2), {"maximum": st.maximum, "material_range": (st.material_min, st.material_max),
                                   "velocity_maximum": st.velocity_maximum}

    def set_tiles(self, family: int, tile_y: int = 8, z_chunk: int = 32) -> None:
        _check(library().seis3d_set_tiles(self._h, family, tile_y, z_chunk))

    def autotune(self, src: Source, steps: int = 5) -> dict:
        """Pick the fastest kernel shape for this grid (cached per grid for the process)."""
        f, y, z = C.c_int(), C.c_int(), C.c_int()
        ms = C.c_double()
        _check(library().seis3d_autotune(self._h, C.byref(src), int(steps), C.byref(f), C.byref(y), C.byref(z),
                                         C.byref(ms)))
        return {"family": f.value, "tile_y": y.value, "z_chunk": z.value, "ms_per_step": ms.value}

    def get_tiles(self) -> tuple[int, int, int]:
        f, y, z = C.c_int(), C.c_int(), C.c_int()
        _check(library().seis3d_get_tiles(self._h, C.byref(f), C.byref(y), C.byref(z)))
        return f.value, y.value, z.value

    # ---- learned absorbing boundary (no reference counterpart) ----

    def abc_attach(self, params=None) -> None:
        """Learned 8-cell frame (pml must be 8); params None gives the plain update (training base)."""
        if params is None:
            _check(library().seis3d_abc_attach(self._h, None))
        else:
            p = np.ascontiguousarray(params, dtype=np.float32)
            if p.size != library().seis3d_abc_param_count():
                raise InvalidArgument("wrong parameter count")
            _check(library().seis3d_abc_attach(self._h, _fptr(p)))

    def abc_detach(self) -> None:
        _check(library().seis3d_abc_detach(self._h))

    def abc_frame_cells(self) -> int:
        n = C.c_int()
        _check(library().seis3d_abc_frame_cells(self._h, C.byref(n)))
        return n.value

    def copy_wavefield_box(self, src: "GPUSimData3D", ox: int, oy: int, oz: int) -> None:
        """This session's non-ghost wavefield from src's padded grid at offset (ox, oy, oz)."""
        _check(library().seis3d_copy_wavefield_box(self._h, src._h, int(ox), int(oy), int(oz)))

    def load_model(self, path) -> None:
        """Stream a .s3dm file of this session's shape into the device model (importModel3D)."""
        _check(library().seis3d_s3dm_load(self._h, os.fsencode(path)))

    def save_model(self, path) -> None:
        """Write the device model to a new .s3dm file (exportModel3D)."""
        _check(library().seis3d_s3dm_save(self._h, os.fsencode(path)))

    def fwi_gradient(self, shots, n_steps: int, stride: int, checkpoint_interval: int = 0):
        """Misfit and gradient ([nz, ny, nx] float32 vp, vs, rho) of shots (source, receivers, observed).

        `observed` of each shot is [n_steps // stride, len(receivers), 3]; shots share this session's
        model; any boundary (CPML, a free top, reflecting walls).
        """
        c = self._ctx_now()
        keep = []
        arr = (Shot * max(1, len(shots)))()
        for k, (src, rec, obs) in enumerate(shots):
            rec = list(rec)
            r = (Receiver * len(rec))(*[Receiver(int(p[0]), int(p[1]), int(p[2])) for p in rec])
            o = np.ascontiguousarray(obs, dtype=np.float32)
            if o.shape != (n_steps // stride, len(rec), 3):
                raise InvalidArgument(f"observed data must be {(n_steps // stride, len(rec), 3)}, got {o.shape}")
            keep += [r, o]
            arr[k] = Shot(src, r, len(rec), _fptr(o))
        shape = (c.nz, c.ny, c.nx)
        g = [np.zeros(shape, np.float32) for _ in range(3)]
        m = C.c_double()
        _check(library().seis3d_fwi_gradient(self._h, arr, len(shots), int(n_steps), int(stride),
                                             int(checkpoint_interval), C.byref(m), *[_fptr(a) for a in g]))
        del keep
        return m.value, g[0], g[1], g[2]


class Batch:
    """Several shots captured into one HIP graph (one launch runs every shot's steps).

    Each session needs its model, receivers and an open gather (gather_begin) before the batch is
    created; after launch() + synchronize() every session's gather holds its shot record.
    """

    MODES = {"shot_streams": 0, "shared_launches": 1, "work_queue": 2}

    def __init__(self, sessions, sources, n_steps: int, shared_launches: bool = False, mode: str | None = None):
        """mode: "shot_streams" (a stream per shot, the default), "shared_launches" (groups of up to
        four shots per launch) or "work_queue" (one persistent launch per half-step for all shots);
        shared_launches=True is the same as mode="shared_launches"."""
        self.sessions = list(sessions)
        n = len(self.sessions)
        if mode is None:
            mode = "shared_launches" if shared_launches else "shot_streams"
        if mode not in self.MODES:
            raise ValueError(f"unknown batch mode {mode!r}")
        hs = (C.c_void_p * n)(*[g.handle.value for g in self.sessions])
        srcs = (Source * n)(*sources)
        self._b = C.c_void_p()
        _check(library().seis3d_batch_create_mode(hs, srcs, n, int(n_steps), self.MODES[mode], C.byref(self._b)))

    def launch(self) -> None:
        _check(library().seis3d_batch_launch(self._b))

    def synchronize(self) -> None:
        _check(library().seis3d_batch_synchronize(self._b))

    def close(self) -> None:
        if self._b.value:
            library().seis3d_batch_destroy(self._b)
        self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fwi_plan(ctx: Ctx, n_steps: int, stride: int, checkpoint_interval: int, receivers: int) -> dict:
    """Device bytes per FWI shot and the checkpoint interval used."""
    b, k = C.c_uint64(), C.c_int()
    _check(library().seis3d_fwi_plan(C.byref(ctx), int(n_steps), int(stride), int(checkpoint_interval),
                                     int(receivers), C.byref(b), C.byref(k)))
    return {"shot_bytes": b.value, "interval": k.value}


def _receiver_array(receivers):
    rec = [tuple(int(v) for v in r) for r in receivers]
    arr = (Receiver * max(1, len(rec)))()
    for i, r in enumerate(rec):
        arr[i] = Receiver(*r)
    return arr, len(rec)


def gather_info(ctx: Ctx, src: Source, receivers, shot_id: int, stride: int, samples: int,
                world_origin=(0.0, 0.0), cmp_bin=(25.0, 25.0)) -> GatherInfo:
    """Frozen gather metadata (startGather3D): sample interval, first sample time, XY scalar."""
    arr, n = _receiver_array(receivers)
    origin = (C.c_double * 2)(*world_origin)
    bins = (C.c_float * 2)(*cmp_bin)
    out = GatherInfo()
    _check(library().seis3d_gather_info_init(C.byref(ctx), C.byref(src), arr, n, int(shot_id), int(stride),
                                             int(samples), origin, bins, C.byref(out)))
    return out


def export_gather(info: GatherInfo, receivers, samples: np.ndarray, directory, prefix: str = "shot") -> list:
    """Write <prefix>_<shot:04>_v{x,y,z}.sgy and _geometry.csv (exportGather3D); returns the paths.

    samples is the gather [sample][receiver][vx, vy, vz]."""
    arr, n = _receiver_array(receivers)
    a = np.ascontiguousarray(samples, dtype=np.float32)
    if a.shape != (info.samples, n, 3):
        raise InvalidArgument(f"gather shape {a.shape} != ({info.samples}, {n}, 3)")
    _check(library().seis3d_export_gather(C.byref(info), arr, n, _fptr(a), os.fsencode(directory),
                                          prefix.encode()))
    base = Path(directory) / f"{prefix}_{info.shot_id:04d}"
    return [Path(f"{base}{e}") for e in ("_vx.sgy", "_vy.sgy", "_vz.sgy", "_geometry.csv")]


def inspect_segy(path) -> SegyInfo:
    info = SegyInfo()
    _check(library().seis3d_segy_inspect(os.fsencode(path), C.byref(info)))
    return info


def read_shot_segy(path) -> dict:
    """One single-component shot (readShotSegy3D): samples [trace][sample], receiver coordinates
    [trace][x, y, z-depth] in metres, first-sample times, trace ids, and integer receiver keys that
    are the coordinate ranks along each axis (world metres are never rounded onto grid nodes)."""
    info = inspect_segy(path)
    samples = np.zeros((info.traces, info.samples), np.float32)
    coords = np.zeros((info.traces, 3), np.float64)
    times = np.zeros(info.traces, np.float64)
    ids = np.zeros(info.traces, np.int32)
    _check(library().seis3d_segy_read(os.fsencode(path), _fptr(samples),
                                      coords.ctypes.data_as(C.POINTER(C.c_double)),
                                      times.ctypes.data_as(C.POINTER(C.c_double)),
                                      ids.ctypes.data_as(C.POINTER(C.c_int32))))
    ranks = np.stack([np.searchsorted(np.unique(coords[:, a]), coords[:, a]) for a in range(3)], axis=1)
    units = ["file units", "Pa", "V", "mV", "A", "m", "m/s", "m/s^2", "N", "W"]
    m = info.measurement_unit
    return {"info": info, "samples": samples, "coordinates": coords, "first_times": times, "trace_ids": ids,
            "receivers": ranks.astype(np.int32), "first_sample_time": float(times[0]),
            "component": ("Vx", "Vy", "Vz")[info.component],
            "amplitude_unit": units[m] if 1 <= m <= 9 else units[0]}


def inspect_model(path, ctx: Ctx | None = None) -> Ctx:
    """Header, payload size and every material cell of a .s3dm file (inspectModel3D)."""
    c = (ctx or default_context()).copy()
    _check(library().seis3d_s3dm_inspect(os.fsencode(path), C.byref(c)))
    return c


def write_model(path, ctx: Ctx, vp: np.ndarray, vs: np.ndarray, rho: np.ndarray) -> None:
    """Write host planes [nz][ny][nx] of Vp, Vs, rho as a new .s3dm file."""
    planes = np.ascontiguousarray(np.stack([vp, vs, rho], axis=-1), dtype=np.float32)
    if planes.shape != (ctx.nz, ctx.ny, ctx.nx, 3):
        raise InvalidArgument(f"model shape {planes.shape[:3]} != ({ctx.nz}, {ctx.ny}, {ctx.nx})")
    _check(library().seis3d_s3dm_write(os.fsencode(path), C.byref(ctx), _fptr(planes)))


def make_view_layout(ctx: Ctx) -> ViewLayout:
    """makeViewLayout3D: slices and volume of the display domain (CPML included, ghosts excluded)."""
    out = ViewLayout()
    _check(library().seis3d_make_view_layout(C.byref(ctx), C.byref(out)))
    return out


def device_count() -> int:
    return int(library().seis3d_device_count())


# Reference-named free functions (cuda/Cuda_Seis3D.cuh:6-42).
def initializeGpu3D(ctx: Ctx) -> GPUSimData3D:
    return GPUSimData3D(ctx)


def releaseGpu3D(g: GPUSimData3D) -> None:
    g.release()


def resetGpu3D(g: GPUSimData3D) -> None:
    g.reset()


def refreshCpmlCoefficients3D(g: GPUSimData3D, ctx: Ctx) -> None:
    g.refresh_cpml(ctx)


def runStep3D(g: GPUSimData3D, src: Source, step: int) -> None:
    g.step(src, step)


def setReceivers3D(g: GPUSimData3D, receivers) -> None:
    g.set_receivers(receivers)


def sampleReceivers3D(g: GPUSimData3D):
    return g.sample_receivers()


def probeMaterial3D(g: GPUSimData3D, x: int, y: int, z: int) -> np.ndarray:
    return g.probe(x, y, z)


def uploadModelPlane3D(g: GPUSimData3D, z: int, values: np.ndarray) -> None:
    g.upload_plane(z, values)


def finishModelUpload3D(g: GPUSimData3D) -> None:
    g.finish_upload()


def downloadModelPlane3D(g: GPUSimData3D, z: int) -> np.ndarray:
    return g.download_plane(z)


def exportGather3D(info: GatherInfo, receivers, samples: np.ndarray, directory, prefix: str = "shot") -> list:
    return export_gather(info, receivers, samples, directory, prefix)


def readShotSegy3D(path) -> dict:
    return read_shot_segy(path)


def inspectModel3D(path, ctx: Ctx | None = None) -> Ctx:
    return inspect_model(path, ctx)


def importModel3D(g: GPUSimData3D, path) -> None:
    g.load_model(path)


def exportModel3D(g: GPUSimData3D, path) -> None:
    g.save_model(path)


def makeViewLayout3D(ctx: Ctx) -> ViewLayout:
    return make_view_layout(ctx)


def extractViews3D(g: GPUSimData3D, slices, component: int, material_component: int, volume: bool,
                   include_boundary: bool = False):
    return g.extract_views(slices, component, material_component, volume, include_boundary)
