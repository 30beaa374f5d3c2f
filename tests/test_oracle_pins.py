"""Independent pins of the oracle's material averaging, free-surface closure, boundaries,
model generator and 2D sources.  CPU only.

Each check states a fact the reference's scheme must satisfy, computed here without the oracle's
code, so that a wrong neighbour, a swapped weight, a wrong sign or an arithmetic mean in place of a
harmonic one fails it:

* face buoyancy (cuda/Cuda_Seis3D.cu:240-241): with a stress linear along an axis the discrete
  derivative is exact, so one step from rest gives v = dt * 2ab/(a+b) * slope at every face,
  where a and b are 1/rho of the two cells sharing that face, for a random 1/rho grid;
* shear moduli (shearMu3D, :275-278): with a velocity linear along an axis one step gives each
  shear stress dt * mu_h * slope, where mu_h is the harmonic mean of the four cells around that
  edge (zero when any of them is fluid), and the normal stresses use the cell's own lambda, mu;
* velocity ghosts (velocityGhostKernel3D, :249-272): a velocity field that meets the two
  zero-traction conditions at the surface is continued linearly into the ghost planes;
* the surface wave: a vertical force at a free surface launches a Rayleigh wave whose phase
  speed, measured between two surface receivers, approaches the root of the Rayleigh equation
  for the model's vp/vs;
* reflecting walls: the reference's own lossless case (96^3, centred explosion, 1600 steps)
  keeps 0.877 of its peak velocity-square sum (docs/reviews/SEIS3D_VOLUME_BOUNDARY_REVIEW.md:45,
  tests/Seis3DBoundaryRegression.cpp:15-56);
* generated models: hand-derived interior and exterior points of every procedural model kind
  (generateMaterialsKernel3D, :114-147);
* 2D sources (applyPointSource, CUDA_Kernels/Cuda_Check.cu:567-640): the totals each source kind
  injects into each field in one step from rest.
"""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

from cmakepro_cuda_seissim_v2_amd import seis3d as S
from oracle import pyoracle as O
from cmakepro_cuda_seissim_v2_amd import seis2d as S2
from oracle import pyoracle2d as O2

VX, VY, VZ, SXX, SYY, SZZ, SXY, SYZ, SXZ = range(9)
INV_RHO, LAMBDA, MU = 9, 10, 11


def grid_ctx(n=24, pml=8, free=False):
    c = S.default_context(nx=n, ny=n, nz=n, pml=pml, free_surface=int(free), model_kind=S.UNIFORM,
                          vp=3000.0, vs=1700.0, density=2200.0, max_vp=4500.0)
    return O.validate(c)


def coords(c):
    Z, Y, X = np.meshgrid(np.arange(c.gz), np.arange(c.gy), np.arange(c.gx), indexing="ij")
    return X.astype(np.float64), Y.astype(np.float64), Z.astype(np.float64)


def interior(c, margin=0):
    """Cells outside every CPML slab and the ghost layers (plus a margin)."""
    lo = 4 + c.pml + margin
    top = (4 + margin) if c.free_surface else lo
    return (slice(top, c.gz - lo), slice(lo, c.gy - lo), slice(lo, c.gx - lo))


def random_materials(c, rng, fluid_fraction=0.0):
    shape = (c.gz, c.gy, c.gx)
    vp = rng.uniform(2000.0, 4500.0, shape)
    vs = vp * rng.uniform(0.3, 0.55, shape)
    if fluid_fraction:
        vs[rng.random(shape) < fluid_fraction] = 0.0
    rho = rng.uniform(1500.0, 3000.0, shape)
    return 1.0 / rho, rho * (vp * vp - 2 * vs * vs), rho * vs * vs


def step_from(c, fields, materials=None, precision=64):
    o = O.Oracle3D(c, precision)
    if materials is not None:
        for f, v in zip((INV_RHO, LAMBDA, MU), materials):
            o.set_field(f, v)
    for f, v in fields.items():
        o.set_field(f, v)
    o.step(S.default_source(x=c.nx // 2, y=c.ny // 2, z=c.nz // 2, amplitude=0.0), 0)
    return o


def shifted(a, axis, by=1):
    """a[i + by * e_axis] (axis 0 = x, 1 = y, 2 = z) on the padded grid, wrapped at the edge."""
    return np.roll(a, -by, axis=2 - axis)


@pytest.mark.parametrize("comp", [0, 1, 2])
def test_face_buoyancy_is_harmonic_over_the_shared_face(comp):
    c = grid_ctx()
    rng = np.random.default_rng(100 + comp)
    ib, lam, mu = random_materials(c, rng)
    X, Y, Z = coords(c)
    slope = 3.0e4
    sigma = slope * (X, Y, Z)[comp]
    o = step_from(c, {(SXX, SYY, SZZ)[comp]: sigma}, (ib, lam, mu))
    a, b = ib, shifted(ib, comp)
    expected = c.dt * (2 * a * b / (a + b)) * slope / c.spacing
    sl = interior(c)
    got = o.field(comp)[sl]
    assert np.max(np.abs(got - expected[sl]) / np.abs(expected[sl])) < 1e-12
    # The arithmetic mean, or the face on the other side, would be far off.
    wrong = c.dt * 0.5 * (a + b) * slope / c.spacing
    assert np.max(np.abs(got - wrong[sl]) / np.abs(wrong[sl])) > 1e-3
    other = c.dt * (2 * a * shifted(ib, comp, -1) / (a + shifted(ib, comp, -1))) * slope / c.spacing
    assert np.max(np.abs(got - other[sl]) / np.abs(other[sl])) > 1e-3
    for k in range(3):
        if k != comp:
            assert np.max(np.abs(o.field(k)[sl])) < 1e-12 * np.max(np.abs(got))


# (velocity component, derivative axis, shear field, the edge's two axes)
SHEAR_CASES = [(0, 1, SXY, (0, 1)), (1, 0, SXY, (0, 1)), (1, 2, SYZ, (1, 2)), (2, 1, SYZ, (1, 2)),
               (0, 2, SXZ, (0, 2)), (2, 0, SXZ, (0, 2))]


@pytest.mark.parametrize("comp,axis,field,edge", SHEAR_CASES)
def test_shear_modulus_is_harmonic_over_the_edge(comp, axis, field, edge):
    c = grid_ctx()
    rng = np.random.default_rng(200 + 3 * comp + axis)
    ib, lam, mu = random_materials(c, rng, fluid_fraction=0.05)
    X, Y, Z = coords(c)
    slope = 0.02
    o = step_from(c, {comp: slope * (X, Y, Z)[axis]}, (ib, lam, mu))
    p, q = edge
    m = [mu, shifted(mu, p), shifted(mu, q), shifted(shifted(mu, p), q)]
    solid = np.all([mk > 0 for mk in m], axis=0)
    with np.errstate(divide="ignore"):
        mu_h = np.where(solid, 4.0 / sum(1.0 / np.where(solid, mk, 1.0) for mk in m), 0.0)
    expected = c.dt * mu_h * slope / c.spacing
    sl = interior(c)
    got = o.field(field)[sl]
    scale = np.max(np.abs(expected[sl]))
    assert np.max(np.abs(got - expected[sl])) < 1e-12 * scale
    assert np.count_nonzero(expected[sl] == 0) > 0 and np.all(got[expected[sl] == 0] == 0)
    # A two-cell mean, or the edge on the other side, would not match.
    with np.errstate(invalid="ignore"):
        alt = c.dt * 2 * mu * shifted(mu, p) / (mu + shifted(mu, p)) * slope / c.spacing
    assert np.nanmax(np.abs(got - alt[sl])) > 1e-3 * scale
    flip = [mu, shifted(mu, p, -1), shifted(mu, q, -1), shifted(shifted(mu, p, -1), q, -1)]
    with np.errstate(divide="ignore"):
        other = 4.0 / sum(1.0 / fk for fk in flip)
    assert np.nanmax(np.abs(got - (c.dt * other * slope / c.spacing)[sl])) > 1e-3 * scale
    for f in range(3, 9):
        if f != field:
            assert np.max(np.abs(o.field(f)[sl])) < 1e-12 * scale


@pytest.mark.parametrize("comp", [0, 1, 2])
def test_normal_stresses_use_the_cell_moduli(comp):
    c = grid_ctx()
    rng = np.random.default_rng(300 + comp)
    ib, lam, mu = random_materials(c, rng)
    X, Y, Z = coords(c)
    slope = 0.02
    o = step_from(c, {comp: slope * (X, Y, Z)[comp]}, (ib, lam, mu))
    d = slope / c.spacing
    sl = interior(c)
    for k in range(3):
        expected = c.dt * (lam * d + (2 * mu * d if k == comp else 0.0))
        got = o.field(3 + k)[sl]
        assert np.max(np.abs(got - expected[sl]) / np.abs(expected[sl])) < 1e-12


def test_velocity_ghosts_continue_a_traction_free_field_vx_vy():
    """vz = c h (x + y)(1 + e (z - 4)): extrapolated from the vz samples at depths h/2 and 3h/2
    to the surface, its horizontal slope is c (1 - e/2).  With vx = vy = -c (1 - e/2) h (z - 4)
    the field satisfies dz vx = -dx vz and dz vy = -dy vz at the surface, and the ghost planes
    must continue vx and vy linearly."""
    c = grid_ctx(free=True)
    X, Y, Z = coords(c)
    cc, e = 2.0e-3, 0.3
    A = -cc * (1 - 0.5 * e) * c.spacing
    vz = cc * c.spacing * ((X - 10) + (Y - 7)) * (1 + e * (Z - 4))
    vx = A * (Z - 4)
    o = step_from(c, {VX: vx, VY: vx.copy(), VZ: vz})
    ghost = (slice(0, 4), slice(4, c.gy - 4), slice(4, c.gx - 4))
    for comp in (VX, VY):
        got, want = o.field(comp)[ghost], vx[ghost]
        assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want)), comp
    # Had the closure extrapolated to depth 0 with weights (2, -1), the slope would be c (1 - e).
    wrong = (-cc * (1 - e) * c.spacing * (Z - 4))[ghost]
    assert np.max(np.abs(o.field(VX)[ghost] - wrong)) > 1e-3 * np.max(np.abs(wrong))


def test_velocity_ghosts_continue_a_traction_free_field_vz():
    """vx = a x, vy = b y and vz = B (z - 3.5) with B = -r (a + b), r = lambda/(lambda + 2 mu) of
    the surface cell: dz vz = -r (dx vx + dy vy) holds, and the vz ghosts continue vz linearly,
    for a random lambda, mu at the surface."""
    c = grid_ctx(free=True)
    rng = np.random.default_rng(7)
    ib, lam, mu = random_materials(c, rng)
    X, Y, Z = coords(c)
    a, b = 1.5e-3, -0.7e-3
    r = lam[4] / (lam[4] + 2 * mu[4])
    B = -r * (a + b)
    vz = B[None, :, :] * (Z - 3.5)
    o = step_from(c, {VX: a * X, VY: b * Y, VZ: vz}, (ib, lam, mu))
    cols = (slice(4, c.gy - 4), slice(4, c.gx - 4))
    got = o.field(VZ)[0:4][(slice(None),) + cols]
    want = vz[0:4][(slice(None),) + cols]
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))
    # With the ratio of the cell below, or an even mirror, the ghosts would differ.
    r_below = lam[5] / (lam[5] + 2 * mu[5])
    wrong = (-r_below * (a + b))[None] * (Z - 3.5)
    assert np.max(np.abs(got - wrong[0:4][(slice(None),) + cols])) > 1e-3 * np.max(np.abs(want))


def rayleigh_speed(vp, vs):
    def f(x):
        return (2 - x * x) ** 2 - 4 * math.sqrt(1 - x * x * vs * vs / (vp * vp)) * math.sqrt(1 - x * x)
    return brentq(f, 0.5, 0.99) * vs


def test_rayleigh_wave_phase_speed():
    vp, vs, rho = 3000.0, 1732.0, 2000.0
    c = S.default_context(nx=200, ny=32, nz=40, pml=12, free_surface=1, model_kind=S.UNIFORM, vp=vp, vs=vs,
                          density=rho, max_vp=vp, frequency=15.0)
    c = O.validate(c)
    cr = rayleigh_speed(vp, vs)
    assert abs(cr / vs - 0.9194) < 1e-3          # Poisson's ratio 1/4: the textbook 0.9194
    sx, offsets = 25, (60, 160)
    rec = [(sx + d, c.ny // 2, 0) for d in offsets]
    src = S.default_source(x=sx, y=c.ny // 2, z=2, kind=S.FORCE, amplitude=1e8, direction=(0.0, 0.0, 1.0))
    n = int((1.2 / 15 + offsets[-1] * c.spacing / (0.95 * vs) + 0.12) / c.dt)
    O.set_threads(0)
    with O.Oracle3D(c, 32) as o:
        o.set_receivers(rec)
        g = o.run(src, 0, n, stride=1, max_samples=n)
    t = (np.arange(n) + 1) * c.dt
    vz = g[:, :, 2]

    def windowed(k):
        tc = offsets[k] * c.spacing / cr + 1.2 / 15
        return vz[:, k] * np.exp(-0.5 * ((t - tc) / 0.08) ** 2)

    A, B = np.fft.rfft(windowed(0), 8192), np.fft.rfft(windowed(1), 8192)
    freq = np.fft.rfftfreq(8192, c.dt)
    dx = (offsets[1] - offsets[0]) * c.spacing
    speeds = []
    for f0 in (7.5, 12.5, 20.0):
        i = int(np.argmin(np.abs(freq - f0)))
        ph = np.angle(B[i] * np.conj(A[i]))
        guess = -2 * np.pi * freq[i] * dx / cr
        ph += 2 * np.pi * np.round((guess - ph) / (2 * np.pi))
        speeds.append(-2 * np.pi * freq[i] * dx / ph)
    # 21 cells per wavelength at 7.5 Hz: within 1 % of the Rayleigh root; the second-order surface
    # closure adds dispersion that grows with frequency but stays within 2.5 % at 20 Hz.
    assert abs(speeds[0] / cr - 1) < 0.01, speeds
    assert speeds[0] < speeds[1] < speeds[2] < 1.025 * cr, speeds


def test_reflecting_walls_reproduce_the_published_lossless_ratio():
    """The reference's lossless case: 96^3, reflecting walls, uniform 3000/1700/2200, 20 Hz
    explosion at the centre, late/peak of the physical collocated sum |v|^2 sampled every 200
    steps up to 1600, published as 0.877."""
    c = S.default_context(nx=96, ny=96, nz=96, pml=12, free_surface=0, reflecting=1, model_kind=S.UNIFORM,
                          vp=3000.0, vs=1700.0, density=2200.0, frequency=20.0)
    c = O.validate(c)
    assert c.dt == pytest.approx(1.271e-3, abs=1e-9)   # 1600 steps end at the published 2.0336 s
    src = S.default_source(x=48, y=48, z=48, kind=S.EXPLOSION, amplitude=1e8)
    O.set_threads(0)
    peak = last = 0.0
    with O.Oracle3D(c, 32) as o:
        for block in range(8):
            o.run(src, block * 200, 200)
            v = [o.field(a).astype(np.float32) for a in range(3)]
            vx = np.float32(0.5) * (v[0][4:-4, 4:-4, 4:-4] + v[0][4:-4, 4:-4, 3:-5])
            vy = np.float32(0.5) * (v[1][4:-4, 4:-4, 4:-4] + v[1][4:-4, 3:-5, 4:-4])
            vz = np.float32(0.5) * (v[2][4:-4, 4:-4, 4:-4] + v[2][3:-5, 4:-4, 4:-4])
            e = float(np.sum(vx.astype(np.float64) ** 2 + vy.astype(np.float64) ** 2 + vz.astype(np.float64) ** 2))
            peak, last = max(peak, e), e
    assert abs(last / peak - 0.877) < 1e-3, last / peak


def generated(kind, **kw):
    c = S.default_context(nx=40, ny=30, nz=50, pml=8, model_kind=kind, **kw)
    c = O.validate(c)
    o = O.Oracle3D(c, 32)
    o.finish_upload() if kind == S.IMPORTED else None
    ib = o.field(INV_RHO)
    o.close()
    return c, ib


def density_at(c, ib, x, y, z):
    return 1.0 / ib[z + c.oz, y + c.oy, x + c.ox]


@pytest.mark.parametrize("kind,inside,outside", [
    # TwoLayer: depth z/(nz-1) >= interface 0.5 is the lower layer.
    (S.TWO_LAYER, [(3, 3, 25), (20, 15, 49)], [(3, 3, 24), (20, 15, 0)]),
    # Sphere: centre (nx/2, ny/2, 0.55 nz); radius 0.187 of each axis' length.
    (S.SPHERE, [(20, 15, 27), (20 + 7, 15, 27), (20, 15, 27 + 9)], [(20 + 8, 15, 27), (0, 0, 0), (20, 15, 18)]),
    # ThreeLayer: thirds of the depth, the middle one half-way between the two materials.
    (S.THREE_LAYER, [(5, 5, 49)], [(5, 5, 0)]),
    # DippingLayer: depth >= 0.5 + 0.25 (x/(nx-1) - 1/2); at x = 0 the interface is at depth 0.375.
    (S.DIPPING_LAYER, [(0, 5, 19), (39, 5, 31)], [(0, 5, 18), (39, 5, 30)]),
    # LowVelocityLayer: |depth - 0.5| <= 0.075.
    # depth 21/49 = 0.4286 and 28/49 = 0.5714 are inside, 20/49 and 29/49 outside.
    (S.LOW_VELOCITY_LAYER, [(5, 5, 21), (5, 5, 28)], [(5, 5, 20), (5, 5, 29)]),
    # Fault: the dipping interface thrown down by 0.18 where x/(nx-1) > 1/2: at x = 19 the
    # interface is at depth 0.4968 (z = 25), at x = 20 at 0.6832 (z = 34).
    (S.FAULT, [(19, 5, 25), (20, 5, 34)], [(19, 5, 24), (20, 5, 33)]),
    # SaltDome: ((x - nx/2)^2/nx^2 + (y - ny/2)^2/ny^2)/0.187^2 + ((depth - 0.78)/0.52)^2 < 1.
    (S.SALT_DOME, [(20, 15, 49), (20, 15, 20)], [(20, 15, 11), (0, 15, 49)]),
])
def test_generated_model_geometry(kind, inside, outside):
    c, ib = generated(kind)
    lower = c.layer_density
    upper = c.density
    for p in inside:
        rho = density_at(c, ib, *p)
        if kind == S.THREE_LAYER:
            assert rho == pytest.approx(lower, rel=1e-6), p
        else:
            assert rho == pytest.approx(lower, rel=1e-6), (kind, p)
    for p in outside:
        assert density_at(c, ib, *p) == pytest.approx(upper, rel=1e-6), (kind, p)


def test_three_layer_middle_and_gradient_blend():
    c, ib = generated(S.THREE_LAYER)
    mid = 0.5 * (c.density + c.layer_density)
    assert density_at(c, ib, 7, 7, 24) == pytest.approx(mid, rel=1e-6)   # depth 24/49 in [1/3, 2/3)
    assert density_at(c, ib, 7, 7, 16) == pytest.approx(c.density, rel=1e-6)   # 16/49 < 1/3
    assert density_at(c, ib, 7, 7, 33) == pytest.approx(c.layer_density, rel=1e-6)   # 33/49 >= 2/3
    c, ib = generated(S.GRADIENT)
    for z in (0, 13, 49):
        want = c.density + (z / 49.0) * (c.layer_density - c.density)
        assert density_at(c, ib, 3, 3, z) == pytest.approx(want, rel=1e-6)
    # Padding repeats the nearest physical cell.
    assert ib[0, 0, 0] == ib[c.oz, c.oy, c.ox]
    assert ib[-1, -1, -1] == ib[c.oz + c.nz - 1, c.oy + c.ny - 1, c.ox + c.nx - 1]


# ---- 2D sources -------------------------------------------------------------------------------


NAMES_2D = ("vx1", "vx2", "vz1", "vz2", "sxx1", "sxx2", "szz1", "szz2", "sxz1", "sxz2")


@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4])
def test_2d_source_totals(kind):
    """One Type-1 step from rest with a single source at the wavelet's peak: the totals that
    applyPointSource defines.  Explosive: 0.5 w into each of sxx1, sxx2, szz1, szz2 (so sxx and szz
    each gain w); shear: 0.5 w into each half of sxz; vertical force: w dt / rho into vz1;
    horizontal force: w dt / rho into vx1; rotated force at 30 degrees: sin 30 = 1/2 of the
    horizontal and cos 30 of the vertical one.  w = wavelet * amplitude * 1e7."""
    from tests.test_oracle2d import make_context
    ctx = make_context(1)
    ctx.src_type = kind
    ctx.src_angle = 30.0
    ctx.src_amp = 0.75
    peak = int(np.argmax(ctx.wavelet))
    w = float(ctx.wavelet[peak]) * 0.75 * 1e7
    dt, rho = ctx.dt, float(ctx.rho[0])
    expect = dict.fromkeys(NAMES_2D, 0.0)
    if kind == S2.VERTICAL_FORCE:
        expect["vz1"] = w * dt / rho
    elif kind == S2.HORIZONTAL_FORCE:
        expect["vx1"] = w * dt / rho
    elif kind == S2.EXPLOSIVE:
        for name in ("sxx1", "sxx2", "szz1", "szz2"):
            expect[name] = 0.5 * w
    elif kind == S2.SHEAR:
        expect["sxz1"] = expect["sxz2"] = 0.5 * w
    else:
        expect["vx1"] = 0.5 * w * dt / rho
        expect["vz1"] = math.sqrt(3.0) / 2 * w * dt / rho
    with O2.Oracle2D(ctx, 64) as o:
        o.step(peak, active_sources=-1)
        fields = {name: o.field(k).reshape(-1) for k, name in enumerate(NAMES_2D)}
    k = ctx.src_idx
    for name, values in fields.items():
        total = float(np.sum(values))
        if expect[name]:
            assert values[k] == pytest.approx(expect[name], rel=1e-6), name
            assert total == pytest.approx(expect[name], rel=1e-6), name
        else:
            assert np.max(np.abs(values)) == 0.0, name
