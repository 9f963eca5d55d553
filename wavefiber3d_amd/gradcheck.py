"""Finite-difference checks of the reverse pass (``wf_po_vjp``) on float32 surfaces.

The boundary takes float32 samples, so a displacement is whatever float32
rounding leaves of it: each check compares the central difference of the
FP64 direct-sum loss with the gradient contracted with the displacement
actually applied, each +/- pair exactly centred on the float32 surface (+d is
rounded to float32 and -d is its exact negation; elements where S0 - d is not
a float32 value get d = 0), and extrapolates the ratio FD / prediction over
two step sizes (Richardson) to remove the O(h^2) curvature of the log loss.

* ``fd_normal_ratio``: n_pts seeded points displaced along their normals
  (x -> x + h v n with v ~ N(0, 1), normals and areas fixed).
* ``fd_tangent_ratio``: the whole surface moved along the FP64 tangent dS/dtau
  of a differentiable surface recipe (all seven inputs), leaving out points
  whose lit state (w.n <= 0) differs between the two ends of the step, where
  the PO currents switch on or off.

A ratio of 1 means the gradient predicts the loss change exactly; the
reverse pass's relative error is |ratio - 1|.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .autograd import fiber_surface_torch, log_intensity_loss, po_intensity


def c5_surface(tau, device="cuda"):
    """The adjoint workload's fiber segment: 40 x 30 um ellipse, length 10 um
    (100 rings), arcs 0..1999 of 2210 per ring (N = 200,000), cuticle 5 um at
    tilt tau (degrees, a 0-d float64 tensor), noise seed 0; 7 float64 tensors."""
    s = fiber_surface_torch(tau, r1=40e-6, r2=30e-6, length=10e-6, spacing=0.1e-6, seed=0,
                            device=device)
    numel = s[0].numel() // 100  # 100 rings, ring-major
    keep = (torch.arange(s[0].numel(), device=s[0].device) % numel) < 2000
    return [t[keep] for t in s]


def loss64(solver, surf, incs, grid, target) -> float:
    """sum (log I - log T)^2 with I from the FP64 direct sum."""
    I = solver.solve_fp64_direct(surf, incs, grid)
    return float(((torch.log(I) - torch.log(target.double())) ** 2).sum().item())


def input_gradients(solver, surf32, incs, grid, target, mode, depth):
    """dL/d(x, y, z, nx, ny, nz, area) (7 float64 tensors) through po_intensity."""
    x = [t.clone().requires_grad_(True) for t in surf32]
    L = log_intensity_loss(po_intensity(x, incs, grid, depth, solver, mode=mode), target)
    L.backward()
    return [t.grad.double() for t in x]


def _centred(S0, d):
    """Zero the elements of d where S0 - d is not a float32 value (then +d and
    -d straddle the float32 surface exactly)."""
    for c in range(len(d)):
        m = (S0[c] - d[c]).float().double() != (S0[c] - d[c])
        d[c][m] = 0
    return d


def fd_normal_ratio(solver, surf32, nrm64, incs, grid, target, grads: dict, n_pts=64, seed=2,
                    h=1e-10):
    """{name: Richardson-extrapolated FD / prediction} for n_pts points moved
    along their normals by h v and h/2 v."""
    n = surf32[0].numel()
    rng = np.random.default_rng(seed)
    idx = torch.from_numpy(rng.choice(n, n_pts, replace=False)).to(surf32[0].device)
    v = torch.zeros(n, dtype=torch.float64, device=surf32[0].device)
    v[idx] = torch.from_numpy(rng.standard_normal(n_pts)).to(v.device)
    base = [surf32[c].double() for c in range(3)]
    ratios = {k: [] for k in grads}
    for step in (h, 0.5 * h):
        d = [(base[c] + step * v * nrm64[c]).float().double() - base[c] for c in range(3)]
        d = _centred(base, d)
        sp = [(base[c] + d[c]).float() for c in range(3)] + list(surf32[3:])
        sm = [(base[c] - d[c]).float() for c in range(3)] + list(surf32[3:])
        fd = loss64(solver, sp, incs, grid, target) - loss64(solver, sm, incs, grid, target)
        for k, g in grads.items():
            pred = sum(float((g[c] * 2 * d[c]).sum().item()) for c in range(3))
            ratios[k].append(fd / pred)
    return {k: (4 * r[1] - r[0]) / 3 for k, r in ratios.items()}


def fd_tangent_ratio(solver, surface_fn, tau0: float, incs, grid, target, grads: dict,
                     steps=(0.002, 0.001), dtau=1e-4):
    """{name: Richardson-extrapolated FD / prediction} for the surface moved
    along its FP64 tangent dS/dtau (central difference of the recipe at
    tau0 +- dtau) by the tilt steps (degrees)."""
    dev = "cuda"
    S0 = [t.float().double() for t in surface_fn(torch.tensor(tau0, dtype=torch.float64, device=dev))]
    up = surface_fn(torch.tensor(tau0 + dtau, dtype=torch.float64, device=dev))
    dn = surface_fn(torch.tensor(tau0 - dtau, dtype=torch.float64, device=dev))
    tdir = [(u.detach() - w.detach()) / (2 * dtau) for u, w in zip(up, dn)]
    wdir = torch.tensor([[math.cos(i.theta) * math.cos(i.phi), math.cos(i.theta) * math.sin(i.phi),
                          math.sin(i.theta)] for i in incs], dtype=torch.float64, device=dev)
    ratios = {k: [] for k in grads}
    for hstep in steps:
        d = [(S0[c] + hstep * tdir[c]).float().double() - S0[c] for c in range(7)]
        d = _centred(S0, d)
        sp = [S0[c] + d[c] for c in range(7)]
        sm = [S0[c] - d[c] for c in range(7)]
        flip = ((torch.stack(sp[3:6], 1) @ wdir.T <= 0)
                != (torch.stack(sm[3:6], 1) @ wdir.T <= 0)).any(1)
        for c in range(7):
            d[c][flip] = 0
        fd = (loss64(solver, [(S0[c] + d[c]).float() for c in range(7)], incs, grid, target)
              - loss64(solver, [(S0[c] - d[c]).float() for c in range(7)], incs, grid, target))
        for k, g in grads.items():
            pred = sum(float((g[c] * 2 * d[c]).sum().item()) for c in range(7))
            ratios[k].append(fd / pred)
    return {k: (4 * r[1] - r[0]) / 3 for k, r in ratios.items()}
