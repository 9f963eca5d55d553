"""Mie-series check harness for the PO solver (compareMie/mie.ipynb).

The reference compares the PO intensity of a 10 um sphere (400 nm, IOR
1.55 + 0.1i, incidence along +x) on the equator ring with the Mie series from
the third-party PyMieScatt package (``ps.ScatteringFunction``, version not
pinned by the reference; mie.ipynb cells 2-6).  PyMieScatt is not a dependency
here; this module restates the published algorithm it implements -- the
Bohren & Huffman series (Absorption and Scattering of Light by Small
Particles, 1983, ch. 4 and appendix A "BHMIE"):

* size parameter x = 2 pi a n_med / lambda, relative index m = n / n_med;
* series length n_stop = x + 4 x^(1/3) + 2;
* logarithmic derivative D_n(mx) by downward recurrence
  D_{n-1} = n/(mx) - 1/(D_n + n/(mx)) from n_max = max(n_stop, |mx|) + 15;
* Riccati-Bessel psi_n, chi_n by upward recurrence from psi_{-1} = cos x,
  psi_0 = sin x, chi_{-1} = -sin x, chi_0 = cos x; xi_n = psi_n - i chi_n;
* a_n = ((D_n/m + n/x) psi_n - psi_{n-1}) / ((D_n/m + n/x) xi_n - xi_{n-1}),
  b_n = ((m D_n + n/x) psi_n - psi_{n-1}) / ((m D_n + n/x) xi_n - xi_{n-1});
* angular functions pi_n, tau_n (pi_0 = 0, pi_1 = 1,
  pi_n = ((2n-1) mu pi_{n-1} - n pi_{n-2})/(n-1), tau_n = n mu pi_n - (n+1) pi_{n-1});
* S1 = sum (2n+1)/(n(n+1)) (a_n pi_n + b_n tau_n),
  S2 = sum (2n+1)/(n(n+1)) (a_n tau_n + b_n pi_n).

``scattering_function`` returns what the notebook reads from
``ps.ScatteringFunction`` (theta, SL = |S1|^2, SR = |S2|^2, SU = (SL + SR)/2
on the angle grid minAngle..maxAngle in steps of angularResolution degrees,
both ends included).  The notebook scales every PO curve to the Mie curve's
sum, so the overall normalisation of the Mie curve does not enter the
comparison; ``normalization="t"`` divides each curve by its trapezoidal
integral over the angle in radians.

``ring_comparison`` repeats the notebook's processing of a PO intensity grid:
the ring thetao = Nt/2 (theta = 1e-10, the plane of incidence for incidence
along +x), rolled by Np/2 so that index 0 is backscatter, block-averaged with
``blur`` (the notebook's blurfunction, cell 1), scaled to the Mie curve's sum.
"""
from __future__ import annotations

import numpy as np


def mie_coefficients(m: complex, x: float):
    """a_n, b_n for n = 1 .. n_stop (Bohren & Huffman eqs. 4.88, 4.53)."""
    m = complex(m)
    x = float(x)
    nstop = int(np.floor(x + 4.0 * x ** (1.0 / 3.0) + 2.0))
    mx = m * x
    nmx = int(max(nstop, abs(mx)) + 15)
    D = np.zeros(nmx + 1, dtype=np.complex128)
    for n in range(nmx, 0, -1):
        D[n - 1] = n / mx - 1.0 / (D[n] + n / mx)
    a = np.zeros(nstop, dtype=np.complex128)
    b = np.zeros(nstop, dtype=np.complex128)
    psi0, psi1 = np.cos(x), np.sin(x)  # psi_{n-2}, psi_{n-1} at n = 1
    chi0, chi1 = -np.sin(x), np.cos(x)
    xi1 = psi1 - 1j * chi1
    for n in range(1, nstop + 1):
        psi = (2.0 * n - 1.0) / x * psi1 - psi0
        chi = (2.0 * n - 1.0) / x * chi1 - chi0
        xi = psi - 1j * chi
        ta = D[n] / m + n / x
        tb = m * D[n] + n / x
        a[n - 1] = (ta * psi - psi1) / (ta * xi - xi1)
        b[n - 1] = (tb * psi - psi1) / (tb * xi - xi1)
        psi0, psi1 = psi1, psi
        chi0, chi1 = chi1, chi
        xi1 = psi1 - 1j * chi1
    return a, b


def mie_s1s2(m: complex, x: float, mu):
    """Amplitude functions S1, S2 at mu = cos(scattering angle)."""
    a, b = mie_coefficients(m, x)
    mu = np.atleast_1d(np.asarray(mu, dtype=np.float64))
    S1 = np.zeros(mu.shape, dtype=np.complex128)
    S2 = np.zeros(mu.shape, dtype=np.complex128)
    p0 = np.zeros_like(mu)  # pi_{n-1}
    p1 = np.ones_like(mu)  # pi_n at n = 1
    for n in range(1, a.shape[0] + 1):
        if n > 1:
            p0, p1 = p1, ((2.0 * n - 1.0) * mu * p1 - n * p0) / (n - 1.0)
        tau = n * mu * p1 - (n + 1.0) * p0
        f = (2.0 * n + 1.0) / (n * (n + 1.0))
        S1 += f * (a[n - 1] * p1 + b[n - 1] * tau)
        S2 += f * (a[n - 1] * tau + b[n - 1] * p1)
    return S1, S2


def mie_efficiencies(m: complex, x: float):
    """(Q_ext, Q_sca, Q_back, g) (Bohren & Huffman eqs. 4.61, 4.62, 4.74)."""
    a, b = mie_coefficients(m, x)
    n = np.arange(1, a.shape[0] + 1, dtype=np.float64)
    qext = 2.0 / x ** 2 * np.sum((2 * n + 1) * (a + b).real)
    qsca = 2.0 / x ** 2 * np.sum((2 * n + 1) * (np.abs(a) ** 2 + np.abs(b) ** 2))
    qback = np.abs(np.sum((2 * n + 1) * (-1.0) ** n * (a - b))) ** 2 / x ** 2
    g1 = np.sum(n[:-1] * (n[:-1] + 2) / (n[:-1] + 1)
                * (a[:-1] * np.conj(a[1:]) + b[:-1] * np.conj(b[1:])).real)
    g2 = np.sum((2 * n + 1) / (n * (n + 1)) * (a * np.conj(b)).real)
    g = 4.0 / (qsca * x ** 2) * (g1 + g2)
    return float(qext), float(qsca), float(qback), float(g)


def scattering_function(m: complex, wavelength: float, diameter: float, n_medium: float = 1.0,
                        min_angle: float = 0.0, max_angle: float = 180.0,
                        angular_resolution: float = 0.5, normalization: str | None = None):
    """theta [rad], SL = |S1|^2, SR = |S2|^2, SU = (SL + SR)/2 (wavelength and
    diameter in the same unit; angles in degrees, both ends included)."""
    x = np.pi * diameter * n_medium / wavelength
    steps = int(round(1 + (max_angle - min_angle) / angular_resolution))
    theta = np.linspace(min_angle, max_angle, steps) * np.pi / 180.0
    S1, S2 = mie_s1s2(complex(m) / n_medium, x, np.cos(theta))
    SL = np.abs(S1) ** 2
    SR = np.abs(S2) ** 2
    SU = 0.5 * (SL + SR)
    if normalization in ("t", "total"):
        SL, SR, SU = (v / np.trapezoid(v, theta) for v in (SL, SR, SU))
    elif normalization in ("m", "max"):
        SL, SR, SU = (v / v.max() for v in (SL, SR, SU))
    elif normalization is not None:
        raise ValueError(f"unknown normalization {normalization!r}")
    return theta, SL, SR, SU


def blur(v, gap: int) -> np.ndarray:
    """compareMie/mie.ipynb cell 1 (blurfunction): output i averages input
    i*gap .. (i+1)*gap inclusive (gap + 1 samples); the last output averages the
    tail and wraps to input 0."""
    v = np.asarray(v, dtype=np.float64)
    m = len(v) // gap
    out = np.zeros(m)
    for i in range(m - 1):
        out[i] = v[i * gap:(i + 1) * gap + 1].sum() / (gap + 1)
    s = (m - 1) * gap
    out[m - 1] = (v[s:].sum() + v[0]) / (len(v) - s + 1)
    return out


def ring(intensity_grid) -> np.ndarray:
    """mie.ipynb cells 4-6: row Nt/2 of a [Nt, Np] PO grid rolled by Np/2 (index 0 =
    phi = pi, backscatter for incidence along +x)."""
    g = np.asarray(intensity_grid)
    nt, np_ = g.shape
    return np.roll(g[nt // 2, :].astype(np.float64), np_ // 2)


def ring_comparison(po_grid, mie_curve, gap: int = 4):
    """Blurred Mie and PO ring curves, PO scaled to the Mie sum (mie.ipynb cell 4),
    and summary figures: the log10 ratio PO/Mie (median and 90th percentile of its
    magnitude), the correlation of the log curves and the ratio in the forward
    (|scattering angle| < 10 deg) and specular-reflection (90 +- 40 deg) lobes."""
    po = ring(po_grid)
    n = po.shape[0]
    mie = np.asarray(mie_curve, dtype=np.float64)[:n]
    mb = blur(mie, gap)
    pb = blur(po, gap)
    pb = pb * mb.sum() / pb.sum()
    lr = np.log10(pb / mb)
    ang = -180.0 + 360.0 / n * gap * np.arange(mb.shape[0])  # scattering angle of each block
    fwd = np.abs(ang) < 10
    refl = np.abs(np.abs(ang) - 90.0) < 40
    return dict(mie=mb, po=pb, angle_deg=ang,
                log_ratio_median=float(np.median(np.abs(lr))),
                log_ratio_p90=float(np.quantile(np.abs(lr), 0.9)),
                log_corr=float(np.corrcoef(np.log10(pb), np.log10(mb))[0, 1]),
                forward_ratio=float(pb[fwd].sum() / mb[fwd].sum()),
                reflection_ratio=float(pb[refl].sum() / mb[refl].sum()))
