"""Finite-distance validator (compareBEM3d/ellipsoid_simple.cu:229-261,
:384-589): oracle pins, no GPU.

* Sources on the x axis seen along +x lie at distances exactly D - a and
  D + a, so each term is C exp(-ik(D -+ a))/(D -+ a): a wrong sign of the
  source-observer difference, a dropped 1/r or a dropped removal of the
  currents' far-field factor (D0 = 1 m) fails it.
* Sources on the z axis seen from the equator ring are all at sqrt(D^2 + a^2),
  whatever the azimuth.
* Far away the sum tends to the far-field sum: F(D) D exp(ik(D - D0)) -> F_far
  with a phase error k |x|^2 / (2D), and the intensity times D^2 tends to the
  far-field intensity.
"""
import numpy as np
import pytest

from oracle import oracle

LAM = 500e-9
K = 2 * np.pi / LAM


def _rand_c(n, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal((n, 6)) + 1j * r.standard_normal((n, 6))


def test_collinear_sources_exact_distance():
    a = 2.0 ** -17  # float32-exact
    D = 1e-3
    xyz = (np.array([a, -a]), np.zeros(2), np.zeros(2))
    C = _rand_c(2, 0)
    F = oracle.direct_sum_finite(xyz, C, K, np.array([[1.0, 0.0, 0.0]]), D)[0, 0]
    undo = np.exp(1j * K * 1.0)
    expect = (C[0] * np.exp(-1j * K * (D - a)) / (D - a)
              + C[1] * np.exp(-1j * K * (D + a)) / (D + a)) * undo
    np.testing.assert_allclose(F, expect, rtol=1e-9, atol=0)


def test_axial_sources_equator_ring():
    a = 3.0 * 2.0 ** -18
    D = 2e-3
    xyz = (np.zeros(2), np.zeros(2), np.array([a, -a]))
    C = _rand_c(2, 1)
    phi = 2 * np.pi * np.arange(16) / 16
    dirs = np.stack([np.cos(phi), np.sin(phi), np.zeros_like(phi)], 1)
    F = oracle.direct_sum_finite(xyz, C, K, dirs, D)[0]
    r = np.sqrt(D * D + a * a)
    expect = (C[0] + C[1]) * np.exp(-1j * K * (r - 1.0)) / r
    for d in range(len(phi)):
        np.testing.assert_allclose(F[d], expect, rtol=1e-9, atol=0)


def _ellipsoid(n_keep=1500):
    import os

    import wavefiber3d_amd as wf

    g = os.path.join(os.path.dirname(__file__), "golden")
    surf = wf.read_soa_mesh(os.path.join(g, "ellipsoid_3_3_2_float.binary"),
                            os.path.join(g, "ellipsoid_3_3_2_normal_float.binary"),
                            os.path.join(g, "ellipsoid_3_3_2_area_float.binary")).arrays()
    idx = np.linspace(0, surf[0].shape[0] - 1, n_keep).astype(np.int64)
    return [np.ascontiguousarray(v[idx]) for v in surf]


def test_far_limit_matches_far_field_sum():
    x, y, z = _ellipsoid()[:3]
    C = _rand_c(x.shape[0], 2)
    dirs = oracle.output_dirs(6, 12)
    D = 50.0
    Ff = oracle.direct_sum(np.stack([x, y, z]), C, K, dirs)[0]
    Fd = oracle.direct_sum_finite((x, y, z), C, K, dirs, D)[0]
    # phase error of the plane-wave form: k max|x|^2 / (2 D)
    rmax2 = float(np.max(x.astype(np.float64) ** 2 + y.astype(np.float64) ** 2 + z.astype(np.float64) ** 2))
    bound = K * rmax2 / (2 * D) + 1e-8
    scaled = Fd * D * np.exp(1j * K * (D - 1.0))
    err = np.abs(scaled - Ff).max() / np.abs(C).sum(0).max()
    assert err <= 2 * bound, (err, bound)


@pytest.mark.parametrize("pol", ["M", "E"])
def test_far_limit_intensity_scales_with_distance(pol):
    surf = _ellipsoid(800)
    inc = dict(theta=0.3, phi=0.2, lambda_m=LAM, pol=pol, eta=1.55 + 0.3j, beam_sigma_m=1e-2)
    dirs = oracle.output_dirs(4, 8)
    If, _ = oracle.solve_direct(surf, [inc], 0, 0, dirs=dirs)
    D = 30.0
    Id, _ = oracle.solve_finite(surf, [inc], 0, 0, D, dirs=dirs)
    np.testing.assert_allclose(Id[0] * D * D, If[0], rtol=1e-4, atol=1e-6 * If[0].max())
    I1, _ = oracle.solve_finite(surf, [inc], 0, 0, 1.0, dirs=dirs)
    # at 1 m the curvature term k|x|^2/2 (a few 1e-5 rad) still shows
    assert not np.allclose(I1[0], If[0], rtol=1e-9, atol=0)
