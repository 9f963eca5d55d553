"""Mie-series harness (compareMie/mie.ipynb; the Bohren & Huffman series that
PyMieScatt implements): pins on the CPU.

* Bohren & Huffman's BHMIE sample case (m = 1.55, radius 0.525 um, 632.8 nm):
  Q_sca = Q_ext = 3.10543, Q_back = 2.92534 as printed in their appendix A.
* Rayleigh limit (x = 0.01): Q_sca = 8/3 x^4 |K|^2, Q_abs = 4 x Im K,
  K = (m^2 - 1)/(m^2 + 2).
* Optical theorem Q_ext = 4/x^2 Re S(0); S1(0) = S2(0); S1(pi) = -S2(pi);
  Q_sca = 1/x^2 int (|S1|^2 + |S2|^2) sin(theta) dtheta.
* No absorption: Q_ext = Q_sca.  Large sphere: Q_ext -> 2.
* The angle grid and blurfunction of the notebook.
"""
import numpy as np
import pytest

from wavefiber3d_amd import mie


def test_bhmie_sample_case():
    x = 2 * np.pi * 0.525 / 0.6328
    qext, qsca, qback, _ = mie.mie_efficiencies(1.55, x)
    assert round(qsca, 5) == 3.10543
    assert round(qext, 5) == 3.10543
    assert round(qback, 5) == 2.92534


@pytest.mark.parametrize("m", [1.5 + 0.1j, 1.33 + 0.0j, 2.0 + 1.0j])
def test_rayleigh_limit(m):
    x = 0.01
    K = (m * m - 1) / (m * m + 2)
    qext, qsca, _, _ = mie.mie_efficiencies(m, x)
    np.testing.assert_allclose(qsca, 8 / 3 * x ** 4 * abs(K) ** 2, rtol=1e-3)
    if m.imag > 0:
        np.testing.assert_allclose(qext - qsca, 4 * x * K.imag, rtol=1e-3)


@pytest.mark.parametrize("m, x", [(1.55 + 0.1j, 12.0), (1.33 + 0.0j, 30.0), (1.55 + 0.1j, 157.08)])
def test_optical_theorem_and_symmetries(m, x):
    S1, S2 = mie.mie_s1s2(m, x, [1.0, -1.0])
    qext, qsca, _, _ = mie.mie_efficiencies(m, x)
    np.testing.assert_allclose(4 / x ** 2 * S1[0].real, qext, rtol=1e-10)
    assert abs(S1[0] - S2[0]) <= 1e-10 * abs(S1[0])
    assert abs(S1[1] + S2[1]) <= 1e-8 * abs(S1[0])
    th = np.linspace(0, np.pi, 40001)
    A, B = mie.mie_s1s2(m, x, np.cos(th))
    integ = np.trapezoid((np.abs(A) ** 2 + np.abs(B) ** 2) * np.sin(th), th) / x ** 2
    np.testing.assert_allclose(integ, qsca, rtol=1e-4)
    if m.imag == 0:
        np.testing.assert_allclose(qext, qsca, rtol=1e-12)


def test_large_sphere_extinction_paradox():
    qext, _, _, _ = mie.mie_efficiencies(1.55 + 0.1j, 2 * np.pi * 10000 / 400)
    assert abs(qext - 2.0) < 0.1


def test_notebook_angle_grid_and_blur():
    th, SL, SR, SU = mie.scattering_function(1.55 + 0.1j, 400, 20000, min_angle=-180,
                                             max_angle=180, angular_resolution=0.25)
    assert th.shape == (1441,)
    assert th[0] == -np.pi and th[720] == 0.0
    # the forward lobe peaks at theta = 0 and the pattern is symmetric in theta
    assert np.argmax(SU) == 720
    np.testing.assert_allclose(SU[:720], SU[721:][::-1], rtol=1e-9)
    np.testing.assert_allclose(SU, 0.5 * (SL + SR))
    v = np.arange(12, dtype=np.float64)
    b = mie.blur(v, 4)
    np.testing.assert_allclose(b, [(0 + 1 + 2 + 3 + 4) / 5, (4 + 5 + 6 + 7 + 8) / 5,
                                   (8 + 9 + 10 + 11 + 0) / 5])
    _, _, _, SUt = mie.scattering_function(1.5, 400, 1000, normalization="t")
    thh = np.linspace(0, np.pi, SUt.shape[0])
    np.testing.assert_allclose(np.trapezoid(SUt, thh), 1.0, rtol=1e-12)


def test_ring_comparison_identity():
    th, SL, _, _ = mie.scattering_function(1.55 + 0.1j, 400, 2000, min_angle=-180, max_angle=180,
                                           angular_resolution=0.25)
    # a PO grid whose ring equals the Mie curve (after the roll) compares exactly
    ring = SL[:1440]
    grid = np.zeros((4, 1440))
    grid[2] = np.roll(ring, -720)
    r = mie.ring_comparison(grid * 7.0, SL)
    assert r["log_ratio_p90"] < 1e-12 and abs(r["log_corr"] - 1) < 1e-12
    assert abs(r["forward_ratio"] - 1) < 1e-12
