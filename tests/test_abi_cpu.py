"""C-ABI library checks that need no GPU: the library loads, exports every
function include/wf_po.h declares, its structs match the ctypes mirror, and
host-side failures return status codes (never abort)."""
import ctypes as ct
import os
import re

import numpy as np
import pytest

import wavefiber3d_amd as wf
from wavefiber3d_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "wf_po.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(wf_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_function():
    L = wf.lib()
    names = _header_functions()
    assert len(names) >= 13
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTS)


def test_abi_version_and_default_opts():
    L = wf.lib()
    assert L.wf_abi_version() == 5
    o = _lib.Opts()
    L.wf_default_opts(ct.byref(o))
    assert o.struct_size == ct.sizeof(_lib.Opts)
    assert o.device == -1 and o.beta == 4.0 and o.solve_streams == 0 and o.box_enlarge == pytest.approx(1.1)


def _no_gpu():
    L = wf.lib()
    h = ct.c_void_p()
    st = L.wf_create(ct.byref(h), None)
    if st == 0:
        L.wf_destroy(h)
        return False
    return True


def test_create_without_device_or_bad_opts_returns_status():
    L = wf.lib()
    h = ct.c_void_p()
    o = _lib.Opts()
    L.wf_default_opts(ct.byref(o))
    o.struct_size = 8  # ABI mismatch is caught before any device call
    assert L.wf_create(ct.byref(h), ct.byref(o)) == 1
    assert b"struct_size" in L.wf_last_error()
    L.wf_default_opts(ct.byref(o))
    o.box_enlarge = 0.5
    assert L.wf_create(ct.byref(h), ct.byref(o)) == 1
    assert L.wf_create(None, None) == 1
    if _no_gpu():
        assert L.wf_create(ct.byref(h), None) == 4  # WF_ERR_HIP, message set
        assert b"device" in L.wf_last_error()


def test_null_handle_paths():
    L = wf.lib()
    g = _lib.Grid(10, 20)
    assert L.wf_solve(None, None, 0, g, None, None, 0, None) == 1
    assert L.wf_set_geometry(None, None, 0, None) == 1
    assert L.wf_get_stats(None, None) == 1
    L.wf_destroy(None)  # no-op
    t = _lib.Tune()
    assert L.wf_get_tune(None, ct.byref(t)) == 1
    assert L.wf_set_tune(None, ct.byref(t)) == 1
    assert L.wf_autotune(None, None, 0, g, None, None) == 1
    assert L.wf_geometry_info(None, None, None, None) == 1
    assert L.wf_batch_add(None, None, 0, None) == 1
    assert L.wf_batch_solve(None, None, 0, g, None, 0, None) == 1
    assert L.wf_batch_info(None, None, None, None, None) == 1
    L.wf_batch_destroy(None)  # no-op


def test_batch_create_arguments():
    L = wf.lib()
    b = ct.c_void_p()
    assert L.wf_batch_create(None, None, 1) == 1
    assert L.wf_batch_create(ct.byref(b), None, 0) == 1
    assert b"n_streams" in L.wf_last_error()
    assert L.wf_batch_create(ct.byref(b), None, 17) == 1
    if _no_gpu():
        assert L.wf_batch_create(ct.byref(b), None, 2) == 4
        assert not b.value


def test_reference_parameter_grids():
    assert wf.reference_lambda(0) == pytest.approx(400e-9)
    assert wf.reference_lambda(24) == pytest.approx(688e-9)
    assert wf.reference_incidence_theta(90) == pytest.approx(1e-10, abs=1e-12)
    assert wf.reference_incidence_phi(90) == pytest.approx(np.pi / 2)


def test_output_directions_match_oracle():
    import oracle

    d = wf.output_directions(11, 22)
    assert np.abs(d - oracle.output_dirs(11, 22)).max() < 1e-15


def test_fibonacci_sphere():
    r = 2e-6
    s = wf.fibonacci_sphere(5000, r)
    assert s.n == 5000
    assert abs(s.area.astype(np.float64).sum() / (4 * np.pi * r * r) - 1) < 1e-5
    p = np.stack([s.x, s.y, s.z]).astype(np.float64)
    q = np.stack([s.nx, s.ny, s.nz]).astype(np.float64)
    assert np.abs(np.linalg.norm(p, axis=0) - r).max() < 1e-12
    assert np.abs(p / r - q).max() < 1e-6


def test_quad_rule_on_cylinder():
    """geometry.h:3-72 on a circular heightfield: areas sum to the lateral area
    of the inscribed polygonal prism, normals point radially outward."""
    from wavefiber3d_amd.geometry import quad_samples_from_heightfield

    r, L, numel, nv = 5e-6, 2e-6, 400, 40
    ph = 2 * np.pi * np.arange(numel) / numel
    hx = np.repeat((r * np.cos(ph))[:, None], nv, 1)
    hy = np.repeat((r * np.sin(ph))[:, None], nv, 1)
    s = quad_samples_from_heightfield(hx, hy, L)
    chord = 2 * r * np.sin(np.pi / numel)
    # every quad is a chord x unitz rectangle (the last ring reuses its own x, y
    # with the next z, geometry.h:25-28, 45-55)
    assert abs(s.area.astype(np.float64).sum() / (numel * chord * L) - 1) < 1e-5
    p = np.stack([s.x, s.y]).astype(np.float64)
    q = np.stack([s.nx, s.ny, s.nz]).astype(np.float64)
    good = s.area > 0
    cosang = np.einsum("cn,cn->n", p[:, good], q[:2, good]) / np.linalg.norm(p[:, good], axis=0)
    assert cosang.min() > 1 - 1e-6


def test_fiber_segment_recipe():
    s = wf.fiber_segment(r1=8e-6, r2=6e-6, length=2e-6, spacing=0.1e-6, seed=0)
    assert s.n > 5000
    assert np.all(np.isfinite(s.x)) and np.all(s.area >= 0)
    q = np.stack([s.nx, s.ny, s.nz]).astype(np.float64)
    assert np.abs(np.linalg.norm(q, axis=0) - 1).max() < 1e-5
    s2 = wf.fiber_segment(r1=8e-6, r2=6e-6, length=2e-6, spacing=0.1e-6, seed=0)
    assert np.array_equal(s.x, s2.x)  # seeded


def test_torch_fiber_generator_matches_numpy_recipe():
    """The differentiable generator reproduces geometry.fiber_segment, and its
    tilt derivative matches a central difference of the generator."""
    import torch

    from wavefiber3d_amd.autograd import fiber_surface_torch

    kw = dict(r1=6e-6, r2=4e-6, length=2e-6, spacing=0.1e-6, seed=3)
    ref = wf.fiber_segment(**kw)
    t = torch.tensor(3.0, dtype=torch.float64, requires_grad=True)
    s = fiber_surface_torch(t, device="cpu", **kw)
    for a, b in zip(s, ref.arrays()):
        assert np.abs(a.detach().numpy() - b.astype(np.float64)).max() <= 2e-7 * max(1.0, np.abs(b).max())
    (s[0] * 1e6).sum().backward()
    h = 1e-4
    sp = fiber_surface_torch(torch.tensor(3.0 + h, dtype=torch.float64), device="cpu", **kw)
    sm = fiber_surface_torch(torch.tensor(3.0 - h, dtype=torch.float64), device="cpu", **kw)
    fd = ((sp[0] - sm[0]).sum() * 1e6 / (2 * h)).item()
    assert abs(t.grad.item() - fd) <= 1e-6 * max(1.0, abs(fd))


def test_fiber_family_sizes():
    """The uneven multi-segment family: radii 20-50 um, lengths 5-40 um,
    20k-400k samples (spacing widened or narrowed from 0.1 um to stay in range)."""
    import math

    from wavefiber3d_amd.geometry import fiber_family_params, fiber_family_segment

    for seed in range(256):
        p = fiber_family_params(seed)
        assert 20e-6 <= p["r1"] <= 50e-6 and p["r2"] == pytest.approx(0.75 * p["r1"])
        assert 5e-6 <= p["length"] <= 40e-6
        t = np.linspace(0, 2 * np.pi, 4097)
        per = np.sum(np.hypot(np.diff(p["r1"] * np.cos(t)), np.diff(p["r2"] * np.sin(t))))
        n = round(per / p["spacing"]) * round(p["length"] / p["spacing"])
        assert 0.98 * 2e4 <= n <= 1.02 * 4e5, (seed, n)
        assert p["spacing"] >= 0.1e-6 or n >= 0.98 * 2e4
    # the generated segment has exactly the predicted numel x numelvert samples
    p = fiber_family_params(3)
    s = fiber_family_segment(3)
    t = np.linspace(0, 2 * np.pi, 4097)
    per = np.sum(np.hypot(np.diff(p["r1"] * np.cos(t)), np.diff(p["r2"] * np.sin(t))))
    assert s.n == round(per / p["spacing"]) * round(p["length"] / p["spacing"])
    assert math.isfinite(float(s.area.sum())) and (s.area > 0).all()
