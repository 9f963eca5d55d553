"""Surface-sample generators and readers (inputs of the PO boundary).

The solver boundary takes SoA float32 positions, outward unit normals and
areas (compareMie/sphere.cu:92-125).  These helpers produce such samples:

* ``fibonacci_sphere`` -- area-weighted Fibonacci lattice on a sphere.
* ``read_soa_mesh`` -- the reference's ``mesh/*.binary`` layout
  ([x...|y...|z...] float32, [nx...|ny...|nz...], [a...]).
* ``quad_samples_from_heightfield`` -- the quad rule of util/geometry.h:3-72
  (centroid of four ring vertices, half the sum of the two triangle areas,
  normalised sum of the two triangle normals), for fiber heightfields.
* ``fiber_segment`` -- a seeded synthetic fiber heightfield (elliptical
  cross-section, sawtooth cuticle scales, Gaussian-correlated height noise)
  turned into samples with that quad rule.
"""
from __future__ import annotations

import ctypes as ct
import math
from dataclasses import dataclass

import numpy as np


@dataclass
class SurfaceSamples:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    nx: np.ndarray
    ny: np.ndarray
    nz: np.ndarray
    area: np.ndarray

    def arrays(self):
        return [self.x, self.y, self.z, self.nx, self.ny, self.nz, self.area]

    @property
    def n(self) -> int:
        return int(self.x.shape[0])

    def bbox(self):
        p = np.stack([self.x, self.y, self.z]).astype(np.float64)
        return p.min(1), p.max(1)


def _f32(*arrs):
    return [np.ascontiguousarray(a, dtype=np.float32) for a in arrs]


def fibonacci_sphere(n: int, radius: float) -> SurfaceSamples:
    """z_i = 1 - (2i+1)/n, phi_i = i pi (3 - sqrt 5), a_i = 4 pi r^2/n, n = x/r."""
    i = np.arange(n, dtype=np.float64)
    zu = 1.0 - (2.0 * i + 1.0) / n
    ph = i * np.pi * (3.0 - np.sqrt(5.0))
    rr = np.sqrt(np.maximum(0.0, 1.0 - zu * zu))
    ux, uy = rr * np.cos(ph), rr * np.sin(ph)
    a = np.full(n, 4.0 * np.pi * radius * radius / n)
    return SurfaceSamples(*_f32(radius * ux, radius * uy, radius * zu, ux, uy, zu, a))


def read_soa_mesh(points_path: str, normals_path: str, area_path: str) -> SurfaceSamples:
    """Reference mesh files (sphere.cu:94-125 / ellipsoid.cu:84-114): float32 SoA."""
    a = np.fromfile(area_path, dtype=np.float32)
    n = a.shape[0]
    p = np.fromfile(points_path, dtype=np.float32).reshape(3, n)
    q = np.fromfile(normals_path, dtype=np.float32).reshape(3, n)
    return SurfaceSamples(*_f32(p[0], p[1], p[2], q[0], q[1], q[2], a))


def quad_samples_from_heightfield(hx: np.ndarray, hy: np.ndarray, period: float,
                                  start: int = 0, end: int | None = None) -> SurfaceSamples:
    """util/geometry.h:78-178 quad rule.

    hx, hy: [numel, numelvert] ring-vertex x/y (the reference file stores
    x[j*numelvert + i] then y likewise, geometry.h:30-31).  Ring i sits at
    z = i*unitz - period/2 with unitz = period/numelvert (treePO.cu:59).  Quads
    j in [start, end] of each ring are emitted (the last ring reuses its own
    vertices as the upper edge, geometry.h:25-28); output order ring-major.
    """
    hx = np.asarray(hx, dtype=np.float64)
    hy = np.asarray(hy, dtype=np.float64)
    numel, numelvert = hx.shape
    if end is None:
        end = numel - 1
    if np.any((hx == 0) & (hy == 0)):
        raise ValueError("heightfield vertex at the origin (geometry.h zero check)")
    unitz = period / numelvert
    i = np.arange(numelvert)
    z = i * unitz - period * 0.5
    znext = (i + 1) * unitz - period * 0.5
    j = np.arange(start, end + 1)
    jn = np.where(j < numel - 1, j + 1, 0)
    iz = np.where(i < numelvert - 1, i + 1, i)
    # vertices [ring i, quad j]
    p1 = np.stack([hx[j][:, i].T, hy[j][:, i].T, np.broadcast_to(z[:, None], (numelvert, j.size))], -1)
    p2 = np.stack([hx[jn][:, i].T, hy[jn][:, i].T, np.broadcast_to(z[:, None], (numelvert, j.size))], -1)
    p3 = np.stack([hx[j][:, iz].T, hy[j][:, iz].T, np.broadcast_to(znext[:, None], (numelvert, j.size))], -1)
    p4 = np.stack([hx[jn][:, iz].T, hy[jn][:, iz].T, np.broadcast_to(znext[:, None], (numelvert, j.size))], -1)
    c = (p1 + p2 + p3 + p4) / 4.0
    a1 = np.cross(p2 - p1, p3 - p1)
    a2 = np.cross(p4 - p2, p3 - p2)
    area = 0.5 * (np.linalg.norm(a1, axis=-1) + np.linalg.norm(a2, axis=-1))
    nv = a1 + a2
    nv /= np.linalg.norm(nv, axis=-1, keepdims=True)
    c = c.reshape(-1, 3)
    nv = nv.reshape(-1, 3)
    return SurfaceSamples(*_f32(c[:, 0], c[:, 1], c[:, 2], nv[:, 0], nv[:, 1], nv[:, 2],
                                area.reshape(-1)))


def read_heightfield(path: str, numel: int, numelvert: int) -> np.ndarray:
    """Fiber heightfield file (treePO.cu:79-88: raw float32, x block then y
    block, vertex (j, i) at j*numelvert + i) through wf_read_heightfield;
    returns float32 [2 * numel * numelvert]."""
    from ._lib import check, lib

    out = np.empty(2 * numel * numelvert, dtype=np.float32)
    check(lib().wf_read_heightfield(str(path).encode(), numel, numelvert,
                                    out.ctypes.data_as(ct.c_void_p)))
    return out


def heightfield_samples(height, numel: int, numelvert: int, period: float,
                        unitz: float | None = None, start: int = 0, end: int | None = None,
                        stream=None):
    """Surface samples of a fiber heightfield on the GPU (wf_heightfield_points;
    util/geometry.h:78-178 set_sourcepoint, arc window [start, end]).

    height: [2 * numel * numelvert] float32 in the file layout (x block, then
    y block), a numpy array (host path; returns SurfaceSamples) or a CUDA
    tensor (device path; returns the 7 float32 CUDA tensors x, y, z, nx, ny,
    nz, area).  unitz defaults to period / numelvert (treePO.cu:59).  A vertex
    at x = y = 0 raises WfError naming the first sample (the reference exits).
    """
    from ._lib import WF_MEM_DEVICE, WF_MEM_HOST, check, lib

    if end is None:
        end = numel - 1
    if unitz is None:
        unitz = period / numelvert
    n = (end - start + 1) * numelvert
    if hasattr(height, "is_cuda") and height.is_cuda:
        import torch

        h = height.detach().to(torch.float32).contiguous()
        if h.numel() != 2 * numel * numelvert:
            raise ValueError("height must hold 2 * numel * numelvert values")
        out = torch.empty((7, n), dtype=torch.float32, device=h.device)
        st = stream if stream is not None else torch.cuda.current_stream(h.device).cuda_stream
        check(lib().wf_heightfield_points(ct.c_void_p(h.data_ptr()), numel, numelvert, period,
                                          unitz, start, end, ct.c_void_p(out.data_ptr()),
                                          WF_MEM_DEVICE, ct.c_void_p(st)))
        return [out[c] for c in range(7)]
    h = np.ascontiguousarray(np.asarray(height, dtype=np.float32).reshape(-1))
    if h.size != 2 * numel * numelvert:
        raise ValueError("height must hold 2 * numel * numelvert values")
    out = np.empty((7, n), dtype=np.float32)
    check(lib().wf_heightfield_points(h.ctypes.data_as(ct.c_void_p), numel, numelvert, period,
                                      unitz, start, end, out.ctypes.data_as(ct.c_void_p),
                                      WF_MEM_HOST, None))
    return SurfaceSamples(*[out[c] for c in range(7)])


def fiber_heightfield(r1: float, r2: float, length: float, spacing: float,
                      cuticle_period: float = 5e-6, tilt_deg: float = 3.0,
                      noise_sigma: float = 0.05e-6, noise_corr: float = 0.5e-6,
                      seed: int = 0, numel: int | None = None):
    """Synthetic fiber heightfield: ellipse r1 x r2, sawtooth cuticle along z
    (h = tan(tilt) * (z mod period)), plus seeded Gaussian-filtered white noise
    (std sigma, correlation length noise_corr) on the (arc, z) grid.
    Returns (hx [numel, numelvert], hy, numelvert)."""
    t = np.linspace(0, 2 * np.pi, 4097)
    per = np.sum(np.hypot(np.diff(r1 * np.cos(t)), np.diff(r2 * np.sin(t))))
    if numel is None:
        numel = int(round(per / spacing))
    numelvert = int(round(length / spacing))
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((numel, numelvert))
    # Gaussian filter on a periodic-in-arc, periodic-in-z grid via FFT
    ka = np.fft.fftfreq(numel, d=per / numel) * 2 * np.pi
    kz = np.fft.fftfreq(numelvert, d=length / numelvert) * 2 * np.pi
    Ka, Kz = np.meshgrid(ka, kz, indexing="ij")
    filt = np.exp(-0.25 * noise_corr ** 2 * (Ka ** 2 + Kz ** 2))
    h = np.real(np.fft.ifft2(np.fft.fft2(w) * filt))
    h *= noise_sigma / max(h.std(), 1e-30)
    phi = 2 * np.pi * np.arange(numel) / numel
    zc = np.arange(numelvert) * (length / numelvert)
    saw = np.tan(np.radians(tilt_deg)) * np.mod(zc, cuticle_period)
    ex, ey = r1 * np.cos(phi), r2 * np.sin(phi)
    nx, ny = r2 * np.cos(phi), r1 * np.sin(phi)
    nn = np.hypot(nx, ny)
    nx, ny = nx / nn, ny / nn
    rad = saw[None, :] + h
    hx = ex[:, None] + nx[:, None] * rad
    hy = ey[:, None] + ny[:, None] * rad
    return hx, hy, numelvert


def fiber_segment(r1: float = 40e-6, r2: float = 30e-6, length: float = 30e-6,
                  spacing: float = 0.1e-6, seed: int = 0, max_points: int | None = None,
                  **kw) -> SurfaceSamples:
    hx, hy, _ = fiber_heightfield(r1, r2, length, spacing, seed=seed, **kw)
    s = quad_samples_from_heightfield(hx, hy, length)
    if max_points is not None and s.n > max_points:
        s = SurfaceSamples(*[a[:max_points] for a in s.arrays()])
    return s


def fiber_family_params(seed: int, n_min: int = 20_000, n_max: int = 400_000) -> dict:
    """Parameters of one segment of the uneven multi-segment workload: the
    fiber recipe of fiber_segment with radius r1 ~ U[20, 50] um (r2 = 0.75 r1,
    the 40 x 30 um aspect), length ~ U[5, 40] um, seeded by `seed`.  The
    spacing is 0.1 um unless that gives more than n_max (or fewer than n_min)
    samples, in which case it is set to give n_max (n_min); the actual count
    is what quad_samples_from_heightfield returns."""
    rng = np.random.default_rng(10_000 + seed)
    r1 = float(rng.uniform(20e-6, 50e-6))
    r2 = 0.75 * r1
    length = float(rng.uniform(5e-6, 40e-6))
    t = np.linspace(0, 2 * np.pi, 4097)
    per = float(np.sum(np.hypot(np.diff(r1 * np.cos(t)), np.diff(r2 * np.sin(t)))))
    area = per * length
    spacing = 0.1e-6
    if area / spacing ** 2 > n_max:
        spacing = math.sqrt(area / n_max)
    elif area / spacing ** 2 < n_min:
        spacing = math.sqrt(area / n_min)
    return dict(r1=r1, r2=r2, length=length, spacing=spacing, seed=seed)


def fiber_family_segment(seed: int, **kw) -> SurfaceSamples:
    """Segment `seed` of the uneven multi-segment workload (fiber_family_params)."""
    return fiber_segment(**fiber_family_params(seed, **kw))
