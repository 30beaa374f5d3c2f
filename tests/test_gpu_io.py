"""Model and shot-record files through live GPU sessions.

* exportModel3D writes exactly the planes downloadModelPlane3D returns. Those are Vp, Vs, rho
  rebuilt from the stored 1/rho, lambda, mu (cuda/Cuda_Seis3D.cu:720-733), so a reload matches
  the source session to a few ulp (not bit for bit), and 80 steps on it agree to rel-L2 1e-5.
* A gather recorded on the GPU, exported to SEG-Y and read back, is bit-identical per component.
"""
import numpy as np
import pytest

from cmakepro_cuda_seissim_v2_amd import seis3d as S
from cmakepro_cuda_seissim_v2_amd import workloads as W

pytestmark = pytest.mark.gpu


def model(nx, ny, nz, seed=1):
    rng = np.random.default_rng(seed)
    z = np.arange(nz, dtype=np.float32)[:, None, None]
    vp = (1800 + 60 * z + rng.uniform(0, 80, (nz, ny, nx))).astype(np.float32)
    vs = (vp / np.float32(1.75)).astype(np.float32)
    rho = (1900 + 10 * z + np.zeros((nz, ny, nx))).astype(np.float32)
    return vp, vs, rho


def test_model_file_round_trip(gpu, tmp_path):
    vp, vs, rho = model(40, 36, 32)
    ctx = W.imported_context(40, 36, 32, max_vp=float(vp.max()), pml=10, free_surface=True)
    src = S.default_source(x=20, y=18, z=8, kind=S.EXPLOSION, amplitude=1e8)
    path = tmp_path / "model.s3dm"
    with S.GPUSimData3D(ctx) as a:
        a.upload_model(vp, vs, rho)
        S.exportModel3D(a, path)
        c = S.inspectModel3D(path, ctx)
        assert (c.nx, c.ny, c.nz, c.free_surface, c.max_vp) == (40, 36, 32, 1, vp.max())
        assert c.dt == ctx.dt
        payload = np.fromfile(path, "<f4", offset=36).reshape(32, 36, 40, 3)
        for z in range(32):
            assert np.array_equal(payload[z], a.download_plane(z))
        with S.GPUSimData3D(c) as b:
            S.importModel3D(b, path)
            for z in (0, 13, 31):
                pa, pb = a.download_plane(z), b.download_plane(z)
                assert np.max(np.abs(pa - pb) / np.abs(pa)) < 1e-6
            for k in range(80):
                a.step(src, k)
                b.step(src, k)
            for f in (S.VX, S.VZ, S.SXX, S.SXZ):
                fa, fb = a.download_field(f), b.download_field(f)
                assert np.linalg.norm(fa - fb) <= 1e-5 * np.linalg.norm(fa)
        wrong = W.imported_context(40, 36, 33, max_vp=float(vp.max()), pml=10)
        with S.GPUSimData3D(wrong) as w, pytest.raises(S.InvalidArgument, match="dimensions"):
            w.load_model(path)


def test_recorded_gather_export(gpu, tmp_path):
    vp, vs, rho = model(48, 40, 32, seed=2)
    ctx = W.imported_context(48, 40, 32, max_vp=float(vp.max()), pml=10, free_surface=True)
    src = S.default_source(x=24, y=20, z=6, kind=S.FORCE, amplitude=1e8, direction=(0.0, 0.0, 1.0))
    rec = [(4 + 4 * i, 10 + 10 * (i % 3), 0) for i in range(10)]
    with S.GPUSimData3D(ctx) as g:
        g.upload_model(vp, vs, rho)
        g.set_receivers(rec)
        g.gather_begin(60, 2)
        g.run(src, 0, 120)
        gather = g.gather_read()
        g.gather_end()
    assert gather.shape == (60, 10, 3) and np.abs(gather).max() > 0
    info = S.gather_info(ctx, src, rec, 12, 2, 60, world_origin=(1000.0, 2000.0), cmp_bin=(20.0, 20.0))
    paths = S.exportGather3D(info, rec, gather, tmp_path, "gpu")
    for comp, p in enumerate(paths[:3]):
        shot = S.readShotSegy3D(p)
        assert np.array_equal(shot["samples"], gather[:, :, comp].T)
        assert shot["info"].shot_id == 12
