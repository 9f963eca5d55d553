"""2D cross-section PO on the GPU (wf_rough2d_solve; compareBEM2d/rough2d.cu)
against the FP64 oracle: the reference's circle and ellipse cross-sections
(both polarisations, the 25-wavelength sweep in one call, host and device
buffers), a tilted elevation, the facet closed form, and argument errors."""
import os

import numpy as np
import pytest

import wavefiber3d_amd as wf
from oracle import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

G = os.path.join(os.path.dirname(__file__), "golden")


def _xs(name):
    return wf.read_crosssection(os.path.join(G, f"{name}_x.txt"), os.path.join(G, f"{name}_y.txt"))


def _check(I, x, y, inc, nphi=3600):
    ref = oracle.rough2d(x, y, inc.theta, inc.phi, inc.lambda_m, inc.pol, inc.eta, nphi=nphi)
    err = np.abs(I - ref).max() / np.abs(ref).max()
    assert err < 1e-11, (inc, err)


@pytest.mark.parametrize("name", ["circle", "ellipse"])
def test_reference_cross_sections_sweep(name):
    x, y = _xs(name)
    incs = [wf.Rough2dIncidence(wf.rough2d.wavelength(l), pol) for pol in "ME" for l in range(25)]
    print(f"case: rough2d {name} n={x.shape[0]} 50 incidences host", flush=True)
    I = wf.rough2d_solve(x, y, incs)
    assert I.shape == (50, 3600)
    for i in (0, 12, 24, 25, 37, 49):
        _check(I[i], x, y, incs[i])
    print(f"case: rough2d {name} device buffers", flush=True)
    Id = wf.rough2d_solve(torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda(), incs)
    torch.cuda.synchronize()
    assert np.array_equal(Id.cpu().numpy(), I)


def test_tilted_elevation_and_azimuth():
    x, y = _xs("circle")
    incs = [wf.Rough2dIncidence(520e-9, "M", 1.4 + 0.02j, theta=0.3, phi=2.5),
            wf.Rough2dIncidence(610e-9, "E", 1.6 + 0.0j, theta=-0.2, phi=0.7)]
    print("case: rough2d circle tilted", flush=True)
    I = wf.rough2d_solve(x, y, incs, n_phi=1000, distance=2.0, radius=5e-6)
    for i, inc in enumerate(incs):
        ref = oracle.rough2d(x, y, inc.theta, inc.phi, inc.lambda_m, inc.pol, inc.eta, nphi=1000,
                             dist=2.0, radius=5e-6)
        assert np.abs(I[i] - ref).max() / np.abs(ref).max() < 1e-11


@pytest.mark.parametrize("pol", ["M", "E"])
def test_facet_specular_closed_form(pol):
    from test_rough2d_cpu import ETA, K, LAM, RADIUS, fresnel_textbook, wedge

    alpha = 2 * np.pi * 300 / 3600
    W = 4e-6
    x, y = wedge(alpha, W, 64)
    print(f"case: rough2d facet pol={pol}", flush=True)
    I = wf.rough2d_solve(x, y, [wf.Rough2dIncidence(LAM, pol, ETA)])[0]
    rs, rp = fresnel_textbook(alpha, ETA)
    R = rs if pol == "M" else rp
    expect = abs(R) ** 2 * W * W * np.cos(alpha) ** 2 * K / (2 * np.pi) / (2 * RADIUS)
    np.testing.assert_allclose(I[600], expect, rtol=1e-6)


def test_errors():
    x, y = _xs("circle")
    inc = [wf.Rough2dIncidence(500e-9)]
    xb = x.copy()
    yb = y.copy()
    xb[7], yb[7] = xb[6], yb[6]  # segment 6 has zero length
    print("case: rough2d zero-length segment", flush=True)
    with pytest.raises(wf.WfError, match="segment 6"):
        wf.rough2d_solve(xb, yb, inc)
    with pytest.raises(wf.WfError, match="theta"):
        wf.rough2d_solve(x, y, [wf.Rough2dIncidence(500e-9, theta=np.pi / 2)])
    with pytest.raises(wf.WfError, match="pol"):
        wf.rough2d_solve(x, y, [wf.Rough2dIncidence(500e-9, pol="X")])
    with pytest.raises(wf.WfError, match="distance"):
        wf.rough2d_solve(x, y, inc, distance=0.0)
