"""The transverse far-field components the tree aggregates (no GPU).

The intensity of a solve is |E''|^2 with E'' = E_perp - z0 s x H_perp
(util/fields.h:172-204).  The tree aggregates a = t.E - z0 p.H and
b = p.E + z0 t.H per solve instead of the six fields, with t, p the unit
vectors of increasing theta (latitude) and phi; these checks pin the identity
and the sign rule of the great-circle extension on random inputs.
"""
import numpy as np

Z0 = np.sqrt(4e-7 * np.pi / 8.8541878e-12)


def _dirs(th, ph):
    s = np.stack([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), np.sin(th)], -1)
    t = np.stack([-np.sin(th) * np.cos(ph), -np.sin(th) * np.sin(ph), np.cos(th)], -1)
    p = np.stack([-np.sin(ph), np.cos(ph), np.zeros_like(ph)], -1)
    return s, t, p


def _six_field_epilogue(E, H, s):
    """E'' = E_perp - z0 s x H_perp, as the six-field epilogue forms it."""
    Ep = E - np.sum(E * s, -1, keepdims=True) * s
    Hp = H - np.sum(H * s, -1, keepdims=True) * s
    return Ep - Z0 * np.cross(s, Hp)


def test_transverse_components_give_the_intensity_and_field():
    rng = np.random.default_rng(7)
    n = 1000
    th = rng.uniform(-np.pi / 2, np.pi / 2, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    E = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    H = (rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))) / Z0
    s, t, p = _dirs(th, ph)
    o = _six_field_epilogue(E, H, s)
    a = np.sum(t * E, -1) - Z0 * np.sum(p * H, -1)
    b = np.sum(p * E, -1) + Z0 * np.sum(t * H, -1)
    I6 = np.sum(np.abs(o) ** 2, -1)
    np.testing.assert_allclose(np.abs(a) ** 2 + np.abs(b) ** 2, I6, rtol=1e-12)
    np.testing.assert_allclose(a[:, None] * t + b[:, None] * p, o, rtol=1e-11, atol=1e-12 * np.abs(o).max())


def test_great_circle_sign_rule():
    """On the great circle of the column pair (phi, phi + pi), parameterized by
    psi (latitude psi at phi for psi <= pi/2, latitude pi - psi at phi + pi
    beyond), s is smooth while t and p change sign on the second half: the
    smooth sequences of the transverse components are a and -a, b and -b."""
    rng = np.random.default_rng(3)
    ph0 = rng.uniform(0, 2 * np.pi)
    psi = np.linspace(-np.pi / 2 + 1e-3, 3 * np.pi / 2 - 1e-3, 4001)
    first = psi <= np.pi / 2
    lat = np.where(first, psi, np.pi - psi)
    ph = np.where(first, ph0, ph0 + np.pi)
    s, t, p = _dirs(lat, ph)
    s_smooth = np.stack([np.cos(psi) * np.cos(ph0), np.cos(psi) * np.sin(ph0), np.sin(psi)], -1)
    np.testing.assert_allclose(s, s_smooth, atol=1e-12)
    sign = np.where(first, 1.0, -1.0)[:, None]
    t_smooth = np.stack([-np.sin(psi) * np.cos(ph0), -np.sin(psi) * np.sin(ph0), np.cos(psi)], -1)
    p_smooth = np.broadcast_to(np.array([-np.sin(ph0), np.cos(ph0), 0.0]), t.shape)
    np.testing.assert_allclose(sign * t, t_smooth, atol=1e-12)
    np.testing.assert_allclose(sign * p, p_smooth, atol=1e-12)
