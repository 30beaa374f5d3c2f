"""2D shot records and shot packages (host-only part of libseis2d; no GPU).

* SHA-256 against Python's hashlib (OpenSSL) on lengths around every padding boundary.
* The record reader against the reference's fixture (tests/Seis2DRecordRegression.cpp:13-29):
  five sample formats, -2 at sample 3, trace delays 0.01 / 0.02 s through the -10 scalar, m/s.
* Export -> import round trip and the refusal of a bad interval without touching an existing
  file (:69-74).
* The reference's package scenario (:75-95): a Unicode directory, the frozen 8x8 model, the
  manifest's escaped model magic, collision refusal, unsafe prefixes, NaN isolation, no staging
  folders left; then an independent check of the package in the way docs/tools/
  verify_shot_package.py describes it (sizes, hashes, SEG-Y headers, model header, wavelet).
"""
import hashlib
import json
import struct

import numpy as np
import pytest

from cmakepro_cuda_seissim_v2_amd import seis2d as S2


@pytest.mark.parametrize("n", [0, 1, 3, 55, 56, 57, 63, 64, 65, 119, 120, 128, 1000, 65536, 100_003])
def test_sha256_matches_hashlib(tmp_path, n):
    data = np.random.default_rng(n).integers(0, 256, n, dtype=np.uint8).tobytes()
    p = tmp_path / "blob"
    p.write_bytes(data)
    assert S2.sha256_file(p) == hashlib.sha256(data).hexdigest()


def test_sha256_known_answers(tmp_path):
    p = tmp_path / "abc"
    p.write_bytes(b"abc")     # FIPS 180-2 appendix B.1
    assert S2.sha256_file(p) == "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"
    p.write_bytes(b"")
    assert S2.sha256_file(p) == "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"


def write_fixture(path, fmt):
    """tests/Seis2DRecordRegression.cpp:13-29: 2 channels x 12 samples, -2 at sample 3."""
    width = 2 if fmt == 3 else 1 if fmt == 8 else 4
    samples, stride = 12, 240 + 12 * width
    b = bytearray(3600 + 2 * stride)
    u16 = lambda at, v: b.__setitem__(slice(at, at + 2), struct.pack(">H", v & 0xFFFF))
    u32 = lambda at, v: b.__setitem__(slice(at, at + 4), struct.pack(">I", v & 0xFFFFFFFF))
    u16(3216, 1000); u16(3220, samples); u16(3224, fmt); u16(3500, 256)
    for ch in range(2):
        at = 3600 + ch * stride
        u16(at + 108, 100 + ch * 100); u16(at + 114, samples); u16(at + 116, 1000); u16(at + 202, 6); u16(at + 214, -10)
        s = at + 240 + 3 * width
        if fmt == 1:
            u32(s, 0xC1200000)
        elif fmt == 5:
            b[s:s + 4] = struct.pack(">f", -2.0)
        elif fmt == 2:
            u32(s, -2)
        elif fmt == 3:
            u16(s, -2)
        else:
            b[s] = 254
    path.write_bytes(bytes(b))


@pytest.mark.parametrize("fmt", [1, 2, 3, 5, 8])
def test_reference_record_fixture(tmp_path, fmt):
    p = tmp_path / f"format-{fmt}.sgy"
    write_fixture(p, fmt)
    r = S2.readRecordSegy2D(p)
    assert r["traces"].shape == (2, 12) and r["traces"][0][3] == -2 and r["units"] == "m/s"
    assert abs(r["first_times"][0] - .01) < 1e-12 and abs(r["first_times"][1] - .02) < 1e-12
    assert r["sample_interval"] == pytest.approx(1e-3)


def test_record_rejections(tmp_path):
    p = tmp_path / "broken.sgy"
    p.write_text("broken")
    with pytest.raises(S2.SolverRuntimeError, match="truncated"):
        S2.readRecordSegy2D(p)
    write_fixture(p, 5)
    p.write_bytes(p.read_bytes()[:-3])
    with pytest.raises(S2.SolverRuntimeError, match="do not match"):
        S2.readRecordSegy2D(p)


def test_export_round_trip_and_bad_interval(tmp_path):
    values = np.array([[0, 1, -2], [3, 0, -1]], np.float32)
    geometry = [S2.TraceMeta(1, 1, 0, 0, 0, 0, 1), S2.TraceMeta(1, 2, 0, 0, 0, 0, 1)]
    path = tmp_path / "export.sgy"
    S2.writeSegyCaptureFile2D(values, geometry, path, .001)
    r = S2.readRecordSegy2D(path)
    assert np.array_equal(r["traces"], values) and r["units"] == "m/s"
    with pytest.raises(S2.InvalidArgument):
        S2.writeSegyCaptureFile2D(values, geometry, path, .1)    # 100000 us does not fit
    assert np.array_equal(S2.readRecordSegy2D(path)["traces"], values)
    raw = path.read_bytes()
    assert struct.unpack_from(">H", raw, 3216)[0] == 1000 and struct.unpack_from(">H", raw, 3220)[0] == 3
    assert struct.unpack_from(">H", raw, 3224)[0] == 5
    t = 3600
    assert struct.unpack_from(">iiii", raw, t + 8) == (1, 1, 0, 0)
    assert struct.unpack_from(">h", raw, t + 70)[0] == 1 and struct.unpack_from(">H", raw, t + 202)[0] == 6


def frozen_model():
    return dict(nx=8, nz=8, npml=0, formulation=1, source_x=0, source_z=0, source_type=0, capture_start_step=0,
                sampling_stride=1, free_surface=0, multiple_sources=0, spacing=10.0, solver_dt=.001, frequency=20.0,
                delay=0.0, source_amplitude=1.0, source_angle=0.0, lambda_=np.full(64, 2e9, np.float32),
                mu=np.full(64, 1e9, np.float32), rho=np.full(64, 2000, np.float32),
                wavelet=np.array([0, 1, 0], np.float32))


def verify_package(directory):
    """Independent check of a package: the properties docs/tools/verify_shot_package.py:9-58 verifies."""
    manifest = json.loads((directory / "manifest.json").read_text(encoding="utf-8"))
    assert manifest["schema"] == "SeisWaveShot2D-1"
    names = set()
    for item in manifest["files"]:
        assert "/" not in item["path"] and item["path"] not in names
        names.add(item["path"])
        data = (directory / item["path"]).read_bytes()
        assert len(data) == item["bytes"] and hashlib.sha256(data).hexdigest() == item["sha256"]
    assert {"record_vx.sgy", "record_vz.sgy", "geometry.csv"} <= names
    n, ns = manifest["traces"], manifest["samples"]
    for c in ("vx", "vz"):
        raw = (directory / f"record_{c}.sgy").read_bytes()
        assert struct.unpack_from(">H", raw, 3216)[0] == manifest["segy_interval_us"]
        assert struct.unpack_from(">H", raw, 3220)[0] == ns and struct.unpack_from(">H", raw, 3224)[0] == 5
        assert len(raw) == 3600 + n * (240 + ns * 4)
    if manifest["model_included"]:
        magic, nx, nz, h = struct.unpack_from("<8sIIf", (directory / "model.s2dm").read_bytes())
        g = manifest["grid"]
        assert magic == b"S2DMOD1\0" and (nx, nz, h) == (g["nx"], g["nz"], g["spacing_m"])
        assert (directory / "model.s2dm").stat().st_size == 20 + nx * nz * 12
        assert (directory / "source_wavelet.f32le").stat().st_size == manifest["solver"]["nt"] * 4
    return manifest


def test_reference_package_scenario(tmp_path):
    values = np.array([[0, 1, -2], [3, 0, -1]], np.float32)
    geometry = [S2.TraceMeta(1, 1, 12345, 0, 0, 0, -1000), S2.TraceMeta(1, 2, 0, 0, 0, 0, -1000)]
    directory = tmp_path / "中文归档"
    kw = dict(prefix="shot", sample_interval=.001, version="regression")
    published = S2.exportShotPackage2D(directory, values, values, geometry, frozen_model(), **kw)
    assert published.name == "shot_0001"
    assert np.array_equal(S2.readRecordSegy2D(published / "record_vz.sgy")["traces"], values)
    assert (published / "model.s2dm").stat().st_size == 20 + 64 * 12 and (published / "manifest.json").exists()
    text = (published / "manifest.json").read_text(encoding="utf-8")
    assert "S2DMOD1\\\\0" in text                                   # escaped as valid JSON text
    manifest = verify_package(published)
    assert manifest["traces"] == 2 and manifest["samples"] == 3 and manifest["segy_interval_us"] == 1000
    assert manifest["solver"]["fast_math_build"] is False and manifest["source"]["numerical_injection_scale"] == 1e7
    rows = (published / "geometry.csv").read_text().splitlines()
    assert rows[0] == "trace,source_x_m,source_z_m,receiver_x_m,receiver_z_m"
    assert rows[1].split(",")[0] == "1" and float(rows[1].split(",")[1]) == 12345 * 0.001   # 17 digits, as written
    csv = (published / "record_vz.csv").read_text().splitlines()
    assert csv[0] == "time_s,trace_1,trace_2" and [float(v) for v in csv[2].split(",")[1:]] == [1.0, 0.0]
    model = np.frombuffer((published / "model.s2dm").read_bytes(), "<f4", offset=20).reshape(3, 64)
    assert np.all(model[0] == 2e9) and np.all(model[1] == 1e9) and np.all(model[2] == 2000)
    before = (published / "record_vx.sgy").read_bytes()
    with pytest.raises(S2.SolverRuntimeError, match="exists"):
        S2.exportShotPackage2D(directory, values, values, geometry, frozen_model(), **kw)
    assert (published / "record_vx.sgy").read_bytes() == before
    for bad in ("../escape", "CON", "bad:name", "trailing.", "", "lpt3.txt"):
        with pytest.raises(S2.InvalidArgument):
            S2.validateExportPrefix2D(bad)
    S2.validateExportPrefix2D("seis2d_shot")
    invalid = values.copy()
    invalid[0, 0] = np.nan
    with pytest.raises(S2.InvalidArgument):
        S2.exportShotPackage2D(directory, invalid, values, geometry, frozen_model(), shot=2, **kw)
    assert not (directory / "shot_0002").exists()
    assert not any(".part-" in p.name for p in directory.iterdir())
    # Without model and CSV the package still verifies.
    lean = S2.exportShotPackage2D(directory, values, values, geometry, frozen_model(), shot=3, include_csv=False,
                                  include_model=False, **kw)
    m = verify_package(lean)
    assert [f["path"] for f in m["files"]] == ["record_vx.sgy", "record_vz.sgy", "geometry.csv"]


def test_trace_metadata_millimetres():
    g = S2.trace_metadata([(4, 30), (7, 30)], (200, 200), 5.0, shot=9)
    assert (g[0].field_record, g[0].trace_in_record, g[0].source_x, g[0].source_y) == (9, 1, 1_000_000, 1_000_000)
    assert (g[1].group_x, g[1].group_y, g[1].coordinate_scalar) == (35_000, 150_000, -1000)
