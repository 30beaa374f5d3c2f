"""CPU checks of the learned CPML profile tables (seis3d_cpml_profile_coefficients; no reference
counterpart).

The library's host tables must equal the oracle's restatement bit for bit.  The restatement itself
is pinned against what fixes it independently:
  * with the reference's own shapes p = {0, 3, 3, 1, 1} it reproduces makeCoefficients3D
    (cuda/Cuda_Seis3D.cu:45-64) - exactly where the layer depths are binary fractions
    (pml = 8, 16), within one float ulp otherwise;
  * interior cells are the identity (a = 0, b = 1, 1/kappa = 1) and the layer cells are the
    reference's;
  * at the outermost cell (depth 1, alpha = 0) the closed forms 1/kappa = 1/p2,
    b = exp(-e^p0 sigma_max dt / p2), a = (b - 1) / p2 hold;
  * at depth 1/2 (a half-cell position of an 8-cell layer) kappa, sigma and alpha take their
    closed forms 1 + (p2-1) 2^-p1, e^p0 sigma_max 2^-p1 and p3 pi f 2^-p4.
"""
import math

import numpy as np
import pytest

from cmakepro_cuda_seissim_v2_amd import seis3d as S
from cmakepro_cuda_seissim_v2_amd import workloads as W
from oracle import pyoracle as O


def contexts(n=40, seed=11):
    rng = np.random.default_rng(seed)
    for _ in range(n):
        yield W.imported_context(int(rng.integers(16, 90)), int(rng.integers(16, 90)), int(rng.integers(16, 90)),
                                 max_vp=float(rng.uniform(1500, 6000)), pml=int(rng.choice([8, 12, 16, 20, 33, 40])),
                                 free_surface=bool(rng.integers(0, 2)), frequency=float(rng.uniform(5, 40)),
                                 spacing=float(rng.choice([5.0, 10.0, 12.5])))


def random_profile(rng):
    return (float(rng.uniform(-2, 2)), float(rng.uniform(1, 6)), float(rng.uniform(1, 8)), float(rng.uniform(0, 3)),
            float(rng.uniform(0.2, 4)))


def test_library_tables_equal_oracle_restatement():
    rng = np.random.default_rng(3)
    for c in contexts():
        p = random_profile(rng)
        for axis in range(3):
            lib = S.cpml_profile_coefficients(c, p, axis)
            assert np.array_equal(lib.view(np.uint32), O.profile_coefficients(c, p, axis).view(np.uint32)), (p, axis)
        # None is the reference profile.
        assert np.array_equal(S.cpml_profile_coefficients(c, None, 0), S.makeCoefficients3D(c, 0))


def test_reference_shapes_reproduce_make_coefficients():
    for c in contexts(seed=12):
        for axis in range(3):
            ref = S.makeCoefficients3D(c, axis).view(np.int32).astype(np.int64)
            got = O.profile_coefficients(c, S.REFERENCE_PROFILE, axis).view(np.int32).astype(np.int64)
            ulps = int(np.max(np.abs(got - ref)))
            if c.pml in (8, 16):
                assert ulps == 0, (c.pml, axis)
            else:
                assert ulps <= 1, (c.pml, axis, ulps)


def test_profile_closed_forms():
    c = W.imported_context(40, 30, 36, max_vp=4000.0, pml=8, free_surface=True, frequency=20.0)
    p = (0.37, 2.5, 4.0, 0.6, 1.7)
    sigma_max = -4.0 * c.max_vp * math.log(1e-6) / (2.0 * c.pml * c.spacing)
    for axis in range(3):
        n = (c.gx, c.gy, c.gz)[axis]
        t = O.profile_coefficients(c, p, axis).reshape(2, n, 3).astype(np.float64)
        ref = S.makeCoefficients3D(c, axis).reshape(2, n, 3)
        layer = ref[:, :, 0] != 0
        # Identity outside the layer, layer cells where the reference has them.
        assert np.array_equal(t[:, :, 0] != 0, layer)
        assert np.all(t[~layer] == np.array([0.0, 1.0, 1.0]))
        # Depth 1 (clamped; the first ghost cell of the far side, alpha = 0 there).
        i = n - 4
        d = (i - (n - 4 - c.pml - 0.5)) / c.pml
        assert d >= 1.0
        b = math.exp(-math.exp(p[0]) * sigma_max / p[2] * c.dt)
        assert t[0, i, 2] == pytest.approx(1.0 / p[2], rel=1e-7)
        assert t[0, i, 1] == pytest.approx(b, rel=1e-6)
        assert t[0, i, 0] == pytest.approx((b - 1.0) / p[2], rel=1e-5)
        # Depth 1/2: the half-cell position n - 4 - pml - 0.5 + pml/2 on the far side.
        j = n - 4 - c.pml // 2 - 1
        d = (j + 0.5 - (n - 4 - c.pml - 0.5)) / c.pml
        assert d == 0.5
        kappa = 1.0 + (p[2] - 1.0) * 2.0 ** -p[1]
        sigma = math.exp(p[0]) * sigma_max * 2.0 ** -p[1]
        alpha = p[3] * math.pi * c.frequency * 2.0 ** -p[4]
        bh = math.exp(-(sigma / kappa + alpha) * c.dt)
        assert t[1, j, 2] == pytest.approx(1.0 / kappa, rel=1e-7)
        assert t[1, j, 1] == pytest.approx(bh, rel=1e-6)
        assert t[1, j, 0] == pytest.approx(sigma * (bh - 1.0) / (kappa * (sigma + kappa * alpha)), rel=1e-5)


def test_profile_validation():
    c = W.imported_context(20, 20, 20, max_vp=3000.0, pml=8)
    bad = [(6.0, 3, 3, 1, 1), (0, 0.5, 3, 1, 1), (0, 9, 3, 1, 1), (0, 3, 0.5, 1, 1), (0, 3, 21, 1, 1),
           (0, 3, 3, -0.1, 1), (0, 3, 3, 11, 1), (0, 3, 3, 1, 0.05), (0, 3, 3, 1, 9), (float("nan"), 3, 3, 1, 1),
           (0, 3, float("inf"), 1, 1)]
    for p in bad:
        with pytest.raises(S.InvalidArgument):
            S.cpml_profile_coefficients(c, p, 0)
    with pytest.raises(S.InvalidArgument):
        S.cpml_profile_coefficients(c, (0, 3, 3, 1), 0)
