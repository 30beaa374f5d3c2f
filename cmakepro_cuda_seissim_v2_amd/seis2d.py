"""ctypes mirror of the 2D solver's C ABI (include/seis2d.h).

Names follow the reference's 2D API (CUDA_Kernels/Cuda_Check.cuh:127-149) and host numerics
(Main_header/SolverNumerics.h, src/Seis2D/Numerics2D.inc): ``SimulationContext`` holds the same
members as the reference struct (Main_header/Common.h:259-287), ``GPUSimData`` owns a device
session, and the free functions keep the reference's names and argument meaning. Errors raise
``InvalidArgument`` (the reference's std::invalid_argument) or ``SolverRuntimeError``.
There is no CPU fallback: without the built HIP library every call raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libseis2d.so"

SPLIT_PML, HYBRID_PML, FREE_SURFACE = 1, 2, 3
VERTICAL_FORCE, HORIZONTAL_FORCE, EXPLOSIVE, SHEAR, ROTATED_FORCE = 0, 1, 2, 3, 4
FIELDS = {"vx1": 0, "vx2": 1, "vz1": 2, "vz2": 3, "sxx1": 4, "sxx2": 5, "szz1": 6, "szz2": 7, "sxz1": 8,
          "sxz2": 9, "vx": 10, "vz": 11, "sxx": 12, "szz": 13, "sxz": 14}
WEIGHTS = (1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0)   # SolverNumerics.h:15-17
MAX_ACTIVE_SOURCES = 200
SOURCE_SCALE = 1e7


class InvalidArgument(ValueError):
    pass


class SolverRuntimeError(RuntimeError):
    pass


class LogicError(RuntimeError):
    pass


class Ctx2D(C.Structure):
    _fields_ = [("nx", C.c_int32), ("nz", C.c_int32), ("total_grid", C.c_int32), ("h", C.c_float), ("dt", C.c_float),
                ("nt", C.c_int32), ("npml", C.c_int32), ("snap_spacenum", C.c_int32), ("flag_type", C.c_int32),
                ("c1_h", C.c_float), ("c2_h", C.c_float), ("c3_h", C.c_float), ("c4_h", C.c_float),
                ("src_idx", C.c_int32), ("src_z_idx", C.c_int32), ("src_angle", C.c_float), ("src_type", C.c_int32),
                ("src_amp", C.c_float), ("num_rcv", C.c_int32), ("up_flag", C.c_int32)]


_FP = C.POINTER(C.c_float)


class Arrays2D(C.Structure):
    _fields_ = [("rho", _FP), ("mu", _FP), ("lambda_", _FP), ("lambda2mu", _FP), ("dp_flat", _FP), ("dx", _FP),
                ("dx_half", _FP), ("dz", _FP), ("dz_half", _FP), ("wavelet", _FP), ("wavelet_len", C.c_int32),
                ("rcv_grid_idx", C.POINTER(C.c_int32))]


class Source2D(C.Structure):
    """GPUSource (Main_header/Common.h:251-256)."""
    _fields_ = [("idx", C.c_int32), ("t", C.c_float), ("f_peak", C.c_float), ("amp", C.c_float)]


_LIB = None


class TraceMeta(C.Structure):
    """seis2d_trace_meta: SeismicIO::TraceMetadata (field record, trace, source/group X and Z, scalar)."""
    _fields_ = [("field_record", C.c_int32), ("trace_in_record", C.c_int32), ("source_x", C.c_int32),
                ("source_y", C.c_int32), ("group_x", C.c_int32), ("group_y", C.c_int32),
                ("coordinate_scalar", C.c_int32)]


class RecordInfo(C.Structure):
    _fields_ = [("traces", C.c_int32), ("samples", C.c_int32), ("sample_interval_us", C.c_int32),
                ("format", C.c_int32), ("velocity_units", C.c_int32)]


class FrozenModel(C.Structure):
    """seis2d_frozen_model: FrozenModel2D, the solver inputs archived with a shot."""
    _fields_ = [("nx", C.c_int32), ("nz", C.c_int32), ("npml", C.c_int32), ("formulation", C.c_int32),
                ("source_x", C.c_int32), ("source_z", C.c_int32), ("source_type", C.c_int32),
                ("capture_start_step", C.c_int32), ("sampling_stride", C.c_int32), ("free_surface", C.c_int32),
                ("multiple_sources", C.c_int32), ("spacing", C.c_float), ("solver_dt", C.c_float),
                ("frequency", C.c_float), ("delay", C.c_float), ("source_amplitude", C.c_float),
                ("source_angle", C.c_float), ("lambda_", C.POINTER(C.c_float)), ("mu", C.POINTER(C.c_float)),
                ("rho", C.POINTER(C.c_float)), ("wavelet", C.POINTER(C.c_float)), ("wavelet_count", C.c_int32)]


class ExportRequest(C.Structure):
    """seis2d_export_request: ShotExportRequest2D."""
    _fields_ = [("directory", C.c_char_p), ("prefix", C.c_char_p), ("version", C.c_char_p), ("device", C.c_char_p),
                ("shot", C.c_int32), ("samples", C.c_int32), ("sample_interval", C.c_float),
                ("include_csv", C.c_int32), ("include_model", C.c_int32)]


def library() -> C.CDLL:
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"HIP 2D solver library missing: {LIB_PATH} (run __graft_entry__.build())")
        lib = C.CDLL(str(LIB_PATH))
        P = C.POINTER
        vp = C.c_void_p
        sig = {
            "seis2d_courant_limit": (C.c_float, []),
            "seis2d_grid_count": (C.c_int, [C.c_int, C.c_int, P(C.c_int64)]),
            "seis2d_boundary_geometry": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int]),
            "seis2d_sampling_stride": (C.c_int, [C.c_double, C.c_double, P(C.c_int)]),
            "seis2d_ricker_wavelet": (C.c_int, [C.c_int, C.c_float, C.c_float, C.c_float, _FP]),
            "seis2d_pml_profiles": (C.c_int, [P(Ctx2D), _FP, _FP, _FP, _FP, _FP, _FP]),
            "seis2d_validate": (C.c_int, [P(Ctx2D), P(Arrays2D)]),
            "seis2d_create": (C.c_int, [P(Ctx2D), P(Arrays2D), P(vp)]),
            "seis2d_destroy": (None, [vp]),
            "seis2d_update_receivers": (C.c_int, [vp, P(C.c_int32), C.c_int]),
            "seis2d_set_sources": (C.c_int, [vp, P(Source2D), C.c_int]),
            "seis2d_step": (C.c_int, [vp, C.c_int, P(Ctx2D), _FP, C.c_int, C.c_int]),
            "seis2d_copy_wavefield": (C.c_int, [vp, _FP]),
            "seis2d_copy_receiver_records": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, _FP, _FP]),
            "seis2d_record_receiver_step": (C.c_int, [vp, C.c_int, _FP, _FP]),
            "seis2d_synchronize": (C.c_int, [vp]),
            "seis2d_run": (C.c_int, [vp, P(Ctx2D), _FP, C.c_int, C.c_int, C.c_int, _FP, _FP]),
            "seis2d_download_field": (C.c_int, [vp, C.c_int, _FP]),
            "seis2d_upload_field": (C.c_int, [vp, C.c_int, _FP]),
            "seis2d_last_error": (C.c_char_p, []),
            "seis2d_write_segy": (C.c_int, [C.c_char_p, _FP, C.c_int, C.c_int, P(TraceMeta), C.c_float]),
            "seis2d_read_segy_info": (C.c_int, [C.c_char_p, P(RecordInfo)]),
            "seis2d_read_segy": (C.c_int, [C.c_char_p, _FP, P(C.c_double)]),
            "seis2d_validate_prefix": (C.c_int, [C.c_char_p]),
            "seis2d_sha256_file": (C.c_int, [C.c_char_p, C.c_char_p]),
            "seis2d_export_shot_package": (C.c_int, [P(ExportRequest), P(FrozenModel), _FP, _FP, C.c_int,
                                                     P(TraceMeta), C.c_char_p, C.c_size_t]),
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
    msg = library().seis2d_last_error().decode(errors="replace")
    if rc == 1:
        raise InvalidArgument(msg)
    if rc == 3:
        raise LogicError(msg)
    raise SolverRuntimeError(msg)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a: np.ndarray):
    return a.ctypes.data_as(_FP) if a is not None else None


# ---- host numerics -------------------------------------------------------------------------------

def courant_limit() -> float:
    return float(library().seis2d_courant_limit())


def checkedGridElementCount(nx: int, nz: int) -> int:
    n = C.c_int64()
    _check(library().seis2d_grid_count(int(nx), int(nz), C.byref(n)))
    return n.value


def validateBoundaryGeometry(nx: int, nz: int, pml: int, free_surface: bool) -> None:
    _check(library().seis2d_boundary_geometry(int(nx), int(nz), int(pml), int(bool(free_surface))))


def getReceiverSamplingStride(dt: float, rate_hz: float) -> int:
    s = C.c_int()
    _check(library().seis2d_sampling_stride(float(dt), float(rate_hz), C.byref(s)))
    return s.value


def generateRickerWavelet(nt: int, dt: float, f0: float, t0: float) -> np.ndarray:
    out = np.zeros(max(1, nt), np.float32)
    _check(library().seis2d_ricker_wavelet(int(nt), float(dt), float(f0), float(t0), _fp(out)))
    return out[:nt]


@dataclass
class SimulationContext:
    """SimulationContext (Main_header/Common.h:259-287); arrays are float32 [nz*nx], row-major [z][x]."""
    NX: int
    NZ: int
    h: float
    dt: float
    nt: int
    npml: int
    flag_type: int = SPLIT_PML
    upFlag: bool = False
    rho: np.ndarray = None
    mu: np.ndarray = None
    lambda_: np.ndarray = None
    lambda2mu: np.ndarray = None
    dp_flat: np.ndarray = None
    dx: np.ndarray = None
    dx_half: np.ndarray = None
    dz: np.ndarray = None
    dz_half: np.ndarray = None
    src_idx: int = 0
    src_z_idx: int = 0
    wavelet: np.ndarray = None
    src_angle: float = 0.0
    src_type: int = VERTICAL_FORCE
    src_amp: float = 1.0
    rcv_grid_idx: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    snap_spacenum: int = 10

    @property
    def total_grid(self) -> int:
        return self.NX * self.NZ

    @property
    def num_rcv(self) -> int:
        return int(len(self.rcv_grid_idx))

    def weights(self):
        return tuple(float(np.float32(w) / np.float32(self.h)) for w in WEIGHTS)

    def c_struct(self) -> Ctx2D:
        c1, c2, c3, c4 = (float(np.float32(np.float32(w) / np.float32(self.h))) for w in WEIGHTS)
        return Ctx2D(self.NX, self.NZ, self.NX * self.NZ, self.h, self.dt, self.nt, self.npml, self.snap_spacenum,
                     self.flag_type, c1, c2, c3, c4, self.src_idx, self.src_z_idx, self.src_angle, self.src_type,
                     self.src_amp, self.num_rcv, int(bool(self.upFlag)))

    def c_arrays(self):
        """(Arrays2D, keep-alive list) pointing at float32/int32 copies of the members."""
        keep = {}
        for name in ("rho", "mu", "lambda_", "lambda2mu", "dx", "dx_half", "dz", "dz_half", "wavelet"):
            v = getattr(self, name)
            keep[name] = None if v is None else _f32(v).reshape(-1)
        keep["dp_flat"] = None if self.dp_flat is None else _f32(self.dp_flat).reshape(-1)
        keep["rcv"] = np.ascontiguousarray(self.rcv_grid_idx, dtype=np.int32).reshape(-1)
        a = Arrays2D()
        for name in ("rho", "mu", "lambda_", "lambda2mu", "dp_flat", "dx", "dx_half", "dz", "dz_half", "wavelet"):
            setattr(a, name, _fp(keep[name]) if keep[name] is not None else None)
        a.wavelet_len = 0 if keep["wavelet"] is None else len(keep["wavelet"])
        a.rcv_grid_idx = keep["rcv"].ctypes.data_as(C.POINTER(C.c_int32)) if len(keep["rcv"]) else None
        return a, keep


def updatePmlDampingArrays(ctx: SimulationContext) -> SimulationContext:
    """Rebuild ctx.dx/dx_half/dz/dz_half from the model (Numerics2D.inc:36-83)."""
    c = ctx.c_struct()
    dx, dxh = np.zeros(ctx.NX, np.float32), np.zeros(ctx.NX, np.float32)
    dz, dzh = np.zeros(ctx.NZ, np.float32), np.zeros(ctx.NZ, np.float32)
    l2m, rho = _f32(ctx.lambda2mu).reshape(-1), _f32(ctx.rho).reshape(-1)
    _check(library().seis2d_pml_profiles(C.byref(c), _fp(l2m), _fp(rho), _fp(dx), _fp(dxh), _fp(dz), _fp(dzh)))
    ctx.dx, ctx.dx_half, ctx.dz, ctx.dz_half = dx, dxh, dz, dzh
    return ctx


def validateSimulationContext(ctx: SimulationContext) -> None:
    a, keep = ctx.c_arrays()
    c = ctx.c_struct()
    _check(library().seis2d_validate(C.byref(c), C.byref(a)))


class GPUSimData:
    """Device state of one 2D simulation (GPUSimData, Cuda_Check.cuh:68-124), opaque here."""

    def __init__(self):
        self._h = C.c_void_p()
        self.flag_type = 0
        self.total_grid = 0
        self.receiverCapacity = 0
        self.NX = self.NZ = 0

    @property
    def live(self) -> bool:
        return bool(self._h.value)

    def close(self) -> None:
        if self._h.value:
            library().seis2d_destroy(self._h)
        self._h = C.c_void_p()
        self.total_grid = 0
        self.receiverCapacity = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def download_field(self, name) -> np.ndarray:
        k = FIELDS[name] if isinstance(name, str) else int(name)
        out = np.zeros(self.total_grid, np.float32)
        _check(library().seis2d_download_field(self._h, k, _fp(out)))
        return out.reshape(self.NZ, self.NX)

    def upload_field(self, name, values: np.ndarray) -> None:
        k = FIELDS[name] if isinstance(name, str) else int(name)
        a = _f32(values).reshape(-1)
        _check(library().seis2d_upload_field(self._h, k, _fp(a)))

    def set_sources(self, sources) -> None:
        """Upload active sources (GPUSource list of (idx, t, f_peak, amp)), as the 2D driver does."""
        arr = (Source2D * max(1, len(sources)))(*[Source2D(int(s[0]), float(s[1]), float(s[2]), float(s[3]))
                                                  for s in sources])
        _check(library().seis2d_set_sources(self._h, arr, len(sources)))

    def run(self, ctx: SimulationContext, first_it: int, n_steps: int, record: bool = True):
        """n_steps steps replayed from a captured graph; returns ([r][n] vx, [r][n] vz) traces."""
        c = ctx.c_struct()
        w = _f32(ctx.wavelet).reshape(-1)
        nr = self.receiverCapacity
        vx = np.zeros((max(1, nr), n_steps), np.float32) if record else None
        vz = np.zeros((max(1, nr), n_steps), np.float32) if record else None
        _check(library().seis2d_run(self._h, C.byref(c), _fp(w), len(w), int(first_it), int(n_steps),
                                    _fp(vx) if record else None, _fp(vz) if record else None))
        return (vx[:nr], vz[:nr]) if record else None

    def synchronize(self) -> None:
        _check(library().seis2d_synchronize(self._h))


# ---- reference-named API (Cuda_Check.cuh:127-149) -------------------------------------------------

def initializeGpuSimulation(gpu: GPUSimData, ctx: SimulationContext) -> GPUSimData:
    a, keep = ctx.c_arrays()
    c = ctx.c_struct()
    gpu.close()
    _check(library().seis2d_create(C.byref(c), C.byref(a), C.byref(gpu._h)))
    gpu.flag_type = ctx.flag_type
    gpu.total_grid = ctx.total_grid
    gpu.NX, gpu.NZ = ctx.NX, ctx.NZ
    gpu.receiverCapacity = ctx.num_rcv
    return gpu


def releaseGpuSimulation(gpu: GPUSimData) -> None:
    gpu.close()


def updateGpuReceivers(gpu: GPUSimData, receiver_indices) -> None:
    idx = np.ascontiguousarray(receiver_indices, dtype=np.int32).reshape(-1)
    ptr = idx.ctypes.data_as(C.POINTER(C.c_int32)) if len(idx) else None
    _check(library().seis2d_update_receivers(gpu._h, ptr, len(idx)))
    gpu.receiverCapacity = len(idx)


def runGpuSimulationStep(gpu: GPUSimData, current_it: int, ctx: SimulationContext, num_active_sources: int = -1):
    c = ctx.c_struct()
    w = _f32(ctx.wavelet).reshape(-1) if ctx.wavelet is not None else np.zeros(0, np.float32)
    _check(library().seis2d_step(gpu._h, int(current_it), C.byref(c), _fp(w) if len(w) else None, len(w),
                                 int(num_active_sources)))


def copyWavefieldToHost(gpu: GPUSimData) -> np.ndarray:
    out = np.zeros(max(1, gpu.total_grid), np.float32)
    _check(library().seis2d_copy_wavefield(gpu._h, _fp(out)))
    return out[: gpu.total_grid].reshape(gpu.NZ, gpu.NX)


def recordReceiverStepOnGpu(gpu: GPUSimData, num_rcv: int):
    vx = np.zeros(max(1, num_rcv), np.float32)
    vz = np.zeros(max(1, num_rcv), np.float32)
    _check(library().seis2d_record_receiver_step(gpu._h, int(num_rcv), _fp(vx), _fp(vz)))
    return vx[:num_rcv], vz[:num_rcv]


def copyReceiverRecordsFromGpu(gpu: GPUSimData, num_rcv: int, nt: int, current_it: int, rec_vx: np.ndarray,
                               rec_vz: np.ndarray) -> None:
    """Write column current_it of the [num_rcv][nt] float32 traces rec_vx / rec_vz in place."""
    if rec_vx.dtype != np.float32 or rec_vz.dtype != np.float32 or not rec_vx.flags.c_contiguous:
        raise InvalidArgument("receiver traces must be contiguous float32 arrays")
    _check(library().seis2d_copy_receiver_records(gpu._h, int(num_rcv), int(nt), int(current_it), _fp(rec_vx),
                                                  _fp(rec_vz)))


# ---- setups ---------------------------------------------------------------------------------------

# ---- shot records and shot packages (src/Seis2D/ShotExport2D.cpp, RecordData2D.cpp) ----

def _meta_array(geometry):
    arr = (TraceMeta * max(1, len(geometry)))()
    for i, g in enumerate(geometry):
        if isinstance(g, TraceMeta):
            arr[i] = g
        else:
            arr[i] = TraceMeta(**g)
    return arr


def trace_metadata(receivers_xz, source_xz, spacing: float, shot: int = 1) -> list:
    """Millimetre trace geometry of a capture (src/Seis2D/Acquisition2D.inc:61-73): scalar -1000,
    Y holds Z."""
    def mm(grid):
        v = float(grid) * spacing * 1000
        if not np.isfinite(v) or abs(v) > 2 ** 31 - 1:
            raise InvalidArgument("Shot coordinate exceeds SEG-Y millimetre range.")
        return int(np.round(v))
    sx, sz = source_xz
    return [TraceMeta(shot, i + 1, mm(sx), mm(sz), mm(x), mm(z), -1000) for i, (x, z) in enumerate(receivers_xz)]


def writeSegyCaptureFile2D(traces, geometry, path, dt: float) -> None:
    """One component [trace][sample] as SEG-Y, IEEE big-endian, m/s (SeismicIO.cpp:296-414)."""
    a = _f32(np.asarray(traces, np.float32))
    if a.ndim != 2 or a.shape[0] != len(geometry):
        raise InvalidArgument("SEG-Y sample count/interval must fit 1..32767 and geometry must match.")
    _check(library().seis2d_write_segy(str(path).encode(), _fp(a), a.shape[0], a.shape[1], _meta_array(geometry),
                                       float(dt)))


def readRecordSegy2D(path) -> dict:
    """Traces [trace][sample], first-sample times, interval and units (RecordData2D.cpp:35-65)."""
    info = RecordInfo()
    _check(library().seis2d_read_segy_info(str(path).encode(), C.byref(info)))
    traces = np.zeros((info.traces, info.samples), np.float32)
    times = np.zeros(info.traces, np.float64)
    _check(library().seis2d_read_segy(str(path).encode(), _fp(traces), times.ctypes.data_as(C.POINTER(C.c_double))))
    return {"traces": traces, "first_times": times, "sample_interval": info.sample_interval_us * 1e-6,
            "format": info.format, "units": "m/s" if info.velocity_units else "", "name": Path(path).name}


def validateExportPrefix2D(prefix: str) -> None:
    _check(library().seis2d_validate_prefix(prefix.encode()))


def sha256_file(path) -> str:
    out = C.create_string_buffer(65)
    _check(library().seis2d_sha256_file(str(path).encode(), out))
    return out.value.decode()


def exportShotPackage2D(directory, vx, vz, geometry, frozen: dict, *, prefix: str = "seis2d_shot", shot: int = 1,
                        sample_interval: float, version: str = "", device: str = "", include_csv: bool = True,
                        include_model: bool = True) -> Path:
    """Publish <directory>/<prefix>_<shot:04> (exportShotPackage2D); vx, vz are [trace][sample].

    `frozen` holds FrozenModel2D's members (nx, nz, npml, formulation, source_x, source_z, source_type,
    capture_start_step, sampling_stride, free_surface, multiple_sources, spacing, solver_dt, frequency,
    delay, source_amplitude, source_angle, lambda_, mu, rho, wavelet)."""
    ax, az = _f32(np.asarray(vx, np.float32)), _f32(np.asarray(vz, np.float32))
    if ax.ndim != 2 or ax.shape != az.shape or ax.shape[0] != len(geometry):
        raise InvalidArgument("Shot components and geometry differ.")
    keep = {}
    m = FrozenModel()
    for k, v in frozen.items():
        if k in ("lambda_", "mu", "rho", "wavelet"):
            keep[k] = _f32(np.asarray(v, np.float32).reshape(-1))
            setattr(m, k, _fp(keep[k]))
        else:
            setattr(m, k, v)
    m.wavelet_count = int(keep["wavelet"].size) if "wavelet" in keep else 0
    req = ExportRequest(str(directory).encode(), prefix.encode(), version.encode(), device.encode(), int(shot),
                        int(ax.shape[1]), float(sample_interval), int(include_csv), int(include_model))
    out = C.create_string_buffer(4096)
    _check(library().seis2d_export_shot_package(C.byref(req), C.byref(m), _fp(ax), _fp(az), ax.shape[0],
                                                _meta_array(geometry), out, 4096))
    return Path(out.value.decode())


def homogeneous_context(nx: int, nz: int, h: float, dt: float, nt: int, npml: int, vp: float, vs: float, rho: float,
                        flag_type: int = SPLIT_PML, up_flag: bool = False, f0: float = 25.0, t0: float | None = None,
                        src_type: int = EXPLOSIVE, src: tuple[int, int] | None = None, receivers=None,
                        src_amp: float = 1.0, pml_profiles: bool = True) -> SimulationContext:
    """A uniform model set up as setupTestContext does (Numerics2D.inc:86-157).

    mu = rho vs^2, lambda2mu = rho vp^2, lambda = lambda2mu - 2 mu (float), PML profiles from the
    model, free-surface modulus dp = (l2m^2 - lam^2)/l2m at row npml, Ricker of f0 delayed by
    t0 (default 1/f0), source at `src` = (x, z) (default (nx/2, nz/4)), receivers as flat indices
    (default the reference's 100-channel line at row min(nz/2 + 30, nz - 5)).
    """
    n = nx * nz
    rho32, vs32, vp32 = np.float32(rho), np.float32(vs), np.float32(vp)
    mu = rho32 * vs32 * vs32
    l2m = rho32 * vp32 * vp32
    lam = np.float32(l2m - np.float32(2.0) * mu)
    ctx = SimulationContext(NX=nx, NZ=nz, h=h, dt=dt, nt=nt, npml=npml, flag_type=flag_type, upFlag=up_flag)
    ctx.rho = np.full(n, rho32, np.float32)
    ctx.mu = np.full(n, mu, np.float32)
    ctx.lambda_ = np.full(n, lam, np.float32)
    ctx.lambda2mu = np.full(n, l2m, np.float32)
    if pml_profiles:
        updatePmlDampingArrays(ctx)
    else:
        ctx.dx, ctx.dx_half = np.zeros(nx, np.float32), np.zeros(nx, np.float32)
        ctx.dz, ctx.dz_half = np.zeros(nz, np.float32), np.zeros(nz, np.float32)
    dp = np.float32((l2m * l2m - lam * lam) / l2m)
    ctx.dp_flat = np.full(nx, dp, np.float32)
    sx, sz = src if src is not None else (nx // 2, nz // 4)
    ctx.src_z_idx, ctx.src_idx = sz, sz * nx + sx
    ctx.src_type, ctx.src_amp = src_type, src_amp
    ctx.wavelet = generateRickerWavelet(nt, dt, f0, (1.0 / f0) if t0 is None else t0)
    if receivers is None:
        nr = min(100, nx - 8)
        rz = min(nz // 2 + 30, nz - 5)
        step = max(1, (nx - 8) // nr)
        receivers = [rz * nx + 4 + r * step for r in range(nr)]
    ctx.rcv_grid_idx = np.asarray(receivers, np.int32)
    return ctx


def config1_context(receiver_row: int = 30) -> SimulationContext:
    """400 x 400 (grid includes the 20-cell PML frame), h=5 m, dt=0.5 ms, 2000 steps, vp/vs/rho
    3000/1732/2000, type-1 split PML, 25 Hz explosive Ricker (t0 = 1/f0) at (200, 200), 100
    receivers at row `receiver_row` (npml + 10), x = 4 + 3 r."""
    nx = nz = 400
    rec = [receiver_row * nx + 4 + 3 * r for r in range(100)]
    return homogeneous_context(nx, nz, 5.0, 5e-4, 2000, 20, 3000.0, 1732.0, 2000.0, flag_type=SPLIT_PML,
                               f0=25.0, src_type=EXPLOSIVE, src=(200, 200), receivers=rec)
