"""torch.autograd bridge for fitting geometry to far-field patterns.

``po_intensity`` returns the far-field intensity of a surface given as seven
CUDA tensors (x, y, z, nx, ny, nz, area); its backward calls the solver's
hand-written reverse pass (``wf_po_vjp`` in the C-ABI), so gradients reach
any torch computation that produced the surface -- e.g. a per-point normal
displacement x -> x + delta n, or the cuticle tilt of ``fiber_surface_torch``.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .po import POSolver


class _POIntensity(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, z, nx, ny, nz, area, solver: POSolver, incs, grid, depth, mode):
        surf = [t.detach().to(torch.float32).contiguous() for t in (x, y, z, nx, ny, nz, area)]
        if mode == "direct64":
            I = solver.solve_fp64_direct(surf, incs, grid).to(torch.float32)
        else:
            # the forward in the arithmetic of the reverse pass, so that the loss
            # and its gradient see the same intensities
            solver.set_geometry(surf, depth)
            I = solver.grad_forward(incs, grid, mode=mode)
        ctx.solver, ctx.incs, ctx.grid, ctx.depth, ctx.mode = solver, list(incs), grid, depth, mode
        ctx.save_for_backward(*surf)
        return I

    @staticmethod
    def backward(ctx, gI):
        surf = list(ctx.saved_tensors)
        solver = ctx.solver
        solver.set_geometry(surf, 0 if ctx.mode == "direct64" else ctx.depth)
        gx, gn, ga = solver.vjp(ctx.incs, ctx.grid, gI.contiguous(), mode=ctx.mode)
        return gx[0], gx[1], gx[2], gn[0], gn[1], gn[2], ga, None, None, None, None, None


def po_intensity(surface, incidences, grid, depth: int, solver: POSolver, fp64: bool = False,
                 mode: str | None = None):
    """Differentiable far-field intensity [n_inc, Nt, Np] of ``surface``
    (sequence of 7 CUDA tensors).  mode: "tree64" (default: the tree with FP64
    patterns, forward and backward), "tree32" (FP32 patterns) or "direct64"
    (the FP64 direct sum, the reference gradient; also fp64=True)."""
    x, y, z, nx, ny, nz, area = surface
    if fp64:
        mode = "direct64"
    if mode is None:
        mode = "tree64"
    if mode not in POSolver.GRAD_MODES:
        raise ValueError(f"mode must be one of {sorted(POSolver.GRAD_MODES)}")
    return _POIntensity.apply(x, y, z, nx, ny, nz, area, solver, list(incidences), tuple(grid),
                              depth, mode)


def log_intensity_loss(I, target, eps: float = 0.0):
    """L = sum_g (log I_g - log T_g)^2 (the fitting loss of the adjoint workload)."""
    return ((torch.log(I + eps) - torch.log(target + eps)) ** 2).sum()


def _quad_rule_torch(hx, hy, period: float):
    """util/geometry.h:3-72 quad rule in torch (differentiable in hx, hy)."""
    numel, numelvert = hx.shape
    dev, dt = hx.device, hx.dtype
    unitz = period / numelvert
    i = torch.arange(numelvert, device=dev, dtype=dt)
    z = i * unitz - period * 0.5
    zn = (i + 1) * unitz - period * 0.5
    jn = torch.roll(torch.arange(numel, device=dev), -1)
    iz = torch.clamp(torch.arange(numelvert, device=dev) + 1, max=numelvert - 1)

    def P(xs, ys, zs):
        return torch.stack([xs.T, ys.T, zs[:, None].expand(numelvert, numel)], -1)

    p1 = P(hx, hy, z)
    p2 = P(hx[jn], hy[jn], z)
    p3 = P(hx[:, iz], hy[:, iz], zn)
    p4 = P(hx[jn][:, iz], hy[jn][:, iz], zn)
    c = (p1 + p2 + p3 + p4) / 4.0
    a1 = torch.cross(p2 - p1, p3 - p1, dim=-1)
    a2 = torch.cross(p4 - p2, p3 - p2, dim=-1)
    area = 0.5 * (a1.norm(dim=-1) + a2.norm(dim=-1))
    nv = a1 + a2
    nv = nv / nv.norm(dim=-1, keepdim=True)
    c = c.reshape(-1, 3)
    nv = nv.reshape(-1, 3)
    return [c[:, 0], c[:, 1], c[:, 2], nv[:, 0], nv[:, 1], nv[:, 2], area.reshape(-1)]


def fiber_surface_torch(tilt_deg, r1: float = 40e-6, r2: float = 30e-6, length: float = 10e-6,
                        spacing: float = 0.1e-6, cuticle_period: float = 5e-6,
                        noise_sigma: float = 0.05e-6, noise_corr: float = 0.5e-6, seed: int = 0,
                        max_points: int | None = None, device="cuda", dtype=torch.float64):
    """Synthetic fiber segment (same recipe as geometry.fiber_segment) built in
    torch so that the samples are differentiable w.r.t. the cuticle tilt
    (degrees, a 0-d tensor).  Returns 7 tensors (float64)."""
    t = np.linspace(0, 2 * np.pi, 4097)
    per = np.sum(np.hypot(np.diff(r1 * np.cos(t)), np.diff(r2 * np.sin(t))))
    numel = int(round(per / spacing))
    numelvert = int(round(length / spacing))
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((numel, numelvert))
    ka = np.fft.fftfreq(numel, d=per / numel) * 2 * np.pi
    kz = np.fft.fftfreq(numelvert, d=length / numelvert) * 2 * np.pi
    Ka, Kz = np.meshgrid(ka, kz, indexing="ij")
    h = np.real(np.fft.ifft2(np.fft.fft2(w) * np.exp(-0.25 * noise_corr ** 2 * (Ka ** 2 + Kz ** 2))))
    h *= noise_sigma / max(h.std(), 1e-30)
    phi = 2 * np.pi * np.arange(numel) / numel
    zc = np.arange(numelvert) * (length / numelvert)
    ex, ey = r1 * np.cos(phi), r2 * np.sin(phi)
    nx, ny = r2 * np.cos(phi), r1 * np.sin(phi)
    nn = np.hypot(nx, ny)
    nx, ny = nx / nn, ny / nn
    T = lambda a: torch.as_tensor(a, device=device, dtype=dtype)
    tilt = tilt_deg if torch.is_tensor(tilt_deg) else T(tilt_deg)
    saw = torch.tan(tilt * (math.pi / 180.0)) * T(np.mod(zc, cuticle_period))
    rad = saw[None, :] + T(h)
    hx = T(ex)[:, None] + T(nx)[:, None] * rad
    hy = T(ey)[:, None] + T(ny)[:, None] * rad
    s = _quad_rule_torch(hx, hy, length)
    if max_points is not None:
        s = [a[:max_points] for a in s]
    return s
