"""2D cross-section PO (compareBEM2d/rough2d.cu:237-585): oracle pins, no GPU.

* Flat facet, specular direction: a wedge whose front face (width W, normal
  at angle a to the incoming wave, split into m segments) is lit and whose
  two back faces are in shadow.  In the specular direction every sample is
  in phase and PO gives |E''| = |R| W cos(a) sqrt(k / (2 pi D)) with R the
  textbook Fresnel coefficient r_s (TM, E along z) or r_p (TE), so the
  scattering is |R|^2 W^2 cos^2(a) k / (2 pi) / (2 radius).
* Moving the back vertex (shadowed faces only) changes nothing.
* A face exactly facing the wave takes the reference's normal-incidence
  branch, which reflects E and H with the same r_s(0) and radiates no
  backscatter; tilted by 1e-6 rad it takes the oblique branch and gives the
  facet formula.
* A large smooth circle (k a = 157) backscatters |r(0)|^2 / 4 in the
  reference's normalisation (the 2D scattering width pi a |r|^2 of geometric
  optics), to O(1/(k a)).
"""
import numpy as np
import pytest

from oracle import oracle

LAM = 500e-9
K = 2 * np.pi / LAM
ETA = 1.55 + 0.1j
RADIUS = 10e-6


def fresnel_textbook(t, eta):
    root = np.sqrt(eta * eta - np.sin(t) ** 2)
    rs = (np.cos(t) - root) / (np.cos(t) + root)
    rp = (eta * eta * np.cos(t) - root) / (eta * eta * np.cos(t) + root)
    return rs, rp


def wedge(alpha, W, m, back=0.25):
    n = np.array([np.cos(alpha), np.sin(alpha)])
    t = np.array([-np.sin(alpha), np.cos(alpha)])
    front = [-t * W / 2 + t * W * j / m for j in range(m + 1)]
    pts = np.array(front + [-n * W * back])
    return pts[:, 0].copy(), pts[:, 1].copy()


@pytest.mark.parametrize("pol", ["M", "E"])
@pytest.mark.parametrize("alpha_idx", [150, 300])  # 15 and 30 degrees, specular at 2 alpha
def test_facet_specular_closed_form(pol, alpha_idx):
    nphi = 3600
    alpha = 2 * np.pi * alpha_idx / nphi
    W = 4e-6
    x, y = wedge(alpha, W, 64)
    I = oracle.rough2d(x, y, 0.0, np.pi, LAM, pol, ETA, nphi=nphi)
    rs, rp = fresnel_textbook(alpha, ETA)
    R = rs if pol == "M" else rp
    expect = abs(R) ** 2 * W * W * np.cos(alpha) ** 2 * K / (2 * np.pi) / (2 * RADIUS)
    # the specular direction phi = 2 alpha is grid index 2 alpha_idx
    np.testing.assert_allclose(I[2 * alpha_idx], expect, rtol=1e-6)
    # the other polarisation's coefficient does not fit
    Ro = rp if pol == "M" else rs
    assert abs(abs(Ro) ** 2 / abs(R) ** 2 - 1) > 0.05


def test_shadowed_faces_do_not_contribute():
    alpha = np.deg2rad(20.0)
    a = oracle.rough2d(*wedge(alpha, 3e-6, 40, back=0.25), 0.0, np.pi, LAM, "M", ETA, nphi=720)
    b = oracle.rough2d(*wedge(alpha, 3e-6, 40, back=0.1), 0.0, np.pi, LAM, "M", ETA, nphi=720)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("pol", ["M", "E"])
def test_normal_incidence_branch(pol):
    W = 4e-6
    # the incidence azimuth pi gives w = (-1, 1.2e-16, 0): w.n = -1 exactly for n = (1, 0)
    x0, y0 = wedge(0.0, W, 32)
    I0 = oracle.rough2d(x0, y0, 0.0, np.pi, LAM, pol, ETA, nphi=3600)
    x1, y1 = wedge(1e-6, W, 32)
    I1 = oracle.rough2d(x1, y1, 0.0, np.pi, LAM, pol, ETA, nphi=3600)
    rs0 = fresnel_textbook(0.0, ETA)[0]
    facet = abs(rs0) ** 2 * W * W * K / (2 * np.pi) / (2 * RADIUS)
    np.testing.assert_allclose(I1[0], facet, rtol=1e-5)
    assert I0[0] < 1e-6 * facet


@pytest.mark.parametrize("pol", ["M", "E"])
def test_smooth_circle_backscatter_geometric_optics(pol):
    a = RADIUS
    n = 4000
    ang = 2 * np.pi * (np.arange(n) + 0.3) / n  # no normal exactly facing the wave
    x, y = a * np.cos(ang), a * np.sin(ang)
    I = oracle.rough2d(x, y, 0.0, np.pi, 400e-9, pol, ETA, nphi=3600)
    r0 = fresnel_textbook(0.0, ETA)[0]
    # backscatter is phi = 0 for a wave travelling along -x
    np.testing.assert_allclose(I[0], abs(r0) ** 2 / 4, rtol=0.05)


def test_reference_cross_sections_load():
    import os

    import wavefiber3d_amd as wf

    g = os.path.join(os.path.dirname(__file__), "golden")
    for name, n in (("circle", 1000), ("ellipse", 2000)):
        x, y = wf.read_crosssection(os.path.join(g, f"{name}_x.txt"), os.path.join(g, f"{name}_y.txt"))
        assert x.shape == (n,) and y.shape == (n,) and x.dtype == np.float64
        # closed, counter-clockwise, outward normals (dy, -dx)
        area = 0.5 * np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)
        assert area > 0
        I = oracle.rough2d(x, y, 0.0, np.pi, 400e-9, "M", ETA, nphi=360)
        assert np.all(np.isfinite(I)) and I.max() > 0
