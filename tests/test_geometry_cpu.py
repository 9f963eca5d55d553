"""Geometry front-end (util/geometry.h:78-178): oracle pins and the file
reader, no GPU.

* A circular heightfield has closed-form samples: each quad is a rectangle of
  chord 2R sin(pi/numel) by unitz (also the last ring, whose upper edge reuses
  its own vertices one unitz higher), normal radial at the mid-angle,
  centroid R cos(pi/numel) at the mid-angle and mid-height.
* A vertex at x = y = 0 is reported at the first sample of the (i, j) loop
  that touches it, window included.
* A window [start, end] is the corresponding slice of the full output.
* The float32 restatement agrees with the independent float64 numpy quad rule
  of the synthetic fiber generator.
"""
import os

import numpy as np
import pytest

import wavefiber3d_amd as wf
from oracle import oracle


def _circle(numel, numelvert, R):
    phi = 2 * np.pi * np.arange(numel) / numel
    hx = np.repeat((R * np.cos(phi))[:, None], numelvert, 1)
    hy = np.repeat((R * np.sin(phi))[:, None], numelvert, 1)
    return np.concatenate([hx.astype(np.float32).ravel(), hy.astype(np.float32).ravel()])


def test_heightfield_circle_closed_form():
    numel, numelvert, R, period = 64, 6, 5e-6, 3e-6
    unitz = np.float32(period / numelvert)
    out, bad = oracle.heightfield_points(_circle(numel, numelvert, R), numel, numelvert, period,
                                         unitz)
    assert bad == -1
    x, y, z, nx, ny, nz, area = [a.reshape(numelvert, numel) for a in out]
    d = np.pi / numel
    mid = 2 * np.pi * np.arange(numel) / numel + d
    i = np.arange(numelvert)[:, None]
    zc = i * float(unitz) - period / 2 + float(unitz) / 2
    np.testing.assert_allclose(area, 2 * R * np.sin(d) * float(unitz) * np.ones_like(area), rtol=2e-5)
    np.testing.assert_allclose(nx, np.cos(mid)[None, :] + 0 * i, atol=2e-6)
    np.testing.assert_allclose(ny, np.sin(mid)[None, :] + 0 * i, atol=2e-6)
    np.testing.assert_allclose(nz, 0 * nz, atol=1e-7)
    np.testing.assert_allclose(x, R * np.cos(d) * np.cos(mid)[None, :] + 0 * i, atol=R * 1e-6)
    np.testing.assert_allclose(y, R * np.cos(d) * np.sin(mid)[None, :] + 0 * i, atol=R * 1e-6)
    np.testing.assert_allclose(z, zc + 0 * x, atol=period * 1e-6)


@pytest.mark.parametrize("window, expect", [((0, None), (1, 2)), ((3, None), (1, 3))])
def test_heightfield_zero_vertex_first_sample(window, expect):
    numel, numelvert = 16, 5
    h = _circle(numel, numelvert, 4e-6)
    j0, i0 = 3, 2
    h[j0 * numelvert + i0] = 0.0
    h[numel * numelvert + j0 * numelvert + i0] = 0.0
    start, end = window
    _, bad = oracle.heightfield_points(h, numel, numelvert, 1e-6, np.float32(0.2e-6), start, end)
    # samples touching vertex (j0, i0): (j0, i0), (j0-1, i0), (j0, i0-1), (j0-1, i0-1)
    i_exp, j_exp = expect
    assert bad == i_exp * numel + j_exp


def test_heightfield_window_is_slice():
    hx, hy, nv = wf.geometry.fiber_heightfield(8e-6, 6e-6, 2e-6, 0.1e-6, seed=3)
    numel = hx.shape[0]
    h = np.concatenate([hx.astype(np.float32).ravel(), hy.astype(np.float32).ravel()])
    unitz = np.float32(2e-6 / nv)
    full, _ = oracle.heightfield_points(h, numel, nv, 2e-6, unitz)
    for s, e in ((0, numel - 1), (5, 40), (numel - 7, numel - 1)):
        part, bad = oracle.heightfield_points(h, numel, nv, 2e-6, unitz, s, e)
        assert bad == -1
        for a, b in zip(part, full):
            assert np.array_equal(a.reshape(nv, e - s + 1), b.reshape(nv, numel)[:, s:e + 1])


def test_heightfield_matches_float64_quad_rule():
    hx, hy, nv = wf.geometry.fiber_heightfield(10e-6, 8e-6, 3e-6, 0.1e-6, seed=1)
    hx, hy = hx.astype(np.float32), hy.astype(np.float32)
    numel = hx.shape[0]
    h = np.concatenate([hx.ravel(), hy.ravel()])
    out, bad = oracle.heightfield_points(h, numel, nv, 3e-6, np.float32(3e-6 / nv))
    assert bad == -1
    ref = wf.geometry.quad_samples_from_heightfield(hx, hy, 3e-6).arrays()
    for a, b in zip(out, ref):
        assert np.max(np.abs(a - b)) <= 1e-5 * np.max(np.abs(b))


def test_read_heightfield_roundtrip(tmp_path):
    numel, nv = 12, 7
    h = np.random.default_rng(0).standard_normal(2 * numel * nv).astype(np.float32)
    p = os.path.join(tmp_path, "fiber0.binary")
    h.tofile(p)
    assert np.array_equal(wf.read_heightfield(p, numel, nv), h)
    with pytest.raises(wf.WfError):
        wf.read_heightfield(p, numel + 1, nv)  # short file
    with pytest.raises(wf.WfError):
        wf.read_heightfield(os.path.join(tmp_path, "missing.binary"), numel, nv)
