"""2D cross-section PO (compareBEM2d/rough2d.cu:237-585) on the GPU.

``rough2d_solve`` wraps the C-ABI ``wf_rough2d_solve``: a closed polyline
(counter-clockwise, sample i at vertex i with the length and outward normal of
segment i -> i+1) lit by plane waves, intensity at n_phi azimuths in the
incidence's elevation plane, FP64.  ``read_crosssection`` reads the
reference's text geometry (one value per line, rough2d.cu:364-390) and
``wavelength`` is the reference's sweep 400 + 300/25 l nm (rough2d.cu:467).
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass

import numpy as np


@dataclass
class Rough2dIncidence:
    lambda_m: float
    pol: str = "M"            # 'M' (E along z at theta 0) or 'E'
    eta: complex = 1.55 + 0.1j  # rough2d.cu:465-466
    theta: float = 0.0        # rough2d.cu:464
    phi: float = np.pi        # rough2d.cu:471 (one incidence azimuth)


def wavelength(index: int) -> float:
    """rough2d.cu:467-468 (double arithmetic)."""
    return (400.0 + (700.0 - 400.0) / 25.0 * index) * 1e-9


def read_crosssection(xfile: str, yfile: str):
    """rough2d.cu:364-390: one coordinate per line (metres) -> float64 x, y."""
    def read(path):
        with open(path) as fh:
            return np.array([float(line) for line in fh if line.strip()], dtype=np.float64)

    x, y = read(xfile), read(yfile)
    if x.shape != y.shape:
        raise ValueError(f"{xfile} and {yfile} hold different numbers of values")
    return x, y


def _incs(incidences):
    from ._lib import Rough2dIncidence as CInc

    arr = (CInc * len(incidences))()
    for i, t in enumerate(incidences):
        eta = complex(t.eta)
        arr[i] = CInc(t.theta, t.phi, t.lambda_m, eta.real, eta.imag, t.pol.encode())
    return arr


def rough2d_solve(x, y, incidences, n_phi: int = 3600, distance: float = 1.0,
                  radius: float = 10e-6, stream=None):
    """Scattering [n_inc, n_phi] float64 = |E''|^2 D / (2 radius)
    (rough2d.cu:577).  x, y: float64 numpy arrays (host path, numpy out) or CUDA
    tensors (device path, CUDA tensor out).  A zero-length segment raises."""
    from ._lib import WF_MEM_DEVICE, WF_MEM_HOST, check, lib

    incidences = list(incidences)
    arr = _incs(incidences)
    k = len(incidences)
    if hasattr(x, "is_cuda") and x.is_cuda:
        import torch

        xd = x.detach().to(torch.float64).contiguous()
        yd = y.detach().to(torch.float64).contiguous()
        if xd.shape != yd.shape or xd.dim() != 1:
            raise ValueError("x and y must be 1-D of equal length")
        out = torch.empty((k, n_phi), dtype=torch.float64, device=xd.device)
        st = stream if stream is not None else torch.cuda.current_stream(xd.device).cuda_stream
        check(lib().wf_rough2d_solve(ct.c_void_p(xd.data_ptr()), ct.c_void_p(yd.data_ptr()),
                                     xd.shape[0], arr, k, n_phi, distance, radius,
                                     ct.c_void_p(out.data_ptr()), WF_MEM_DEVICE, ct.c_void_p(st)))
        return out
    xh = np.ascontiguousarray(x, dtype=np.float64)
    yh = np.ascontiguousarray(y, dtype=np.float64)
    if xh.shape != yh.shape or xh.ndim != 1:
        raise ValueError("x and y must be 1-D of equal length")
    out = np.empty((k, n_phi), dtype=np.float64)
    check(lib().wf_rough2d_solve(xh.ctypes.data_as(ct.c_void_p), yh.ctypes.data_as(ct.c_void_p),
                                 xh.shape[0], arr, k, n_phi, distance, radius,
                                 out.ctypes.data_as(ct.c_void_p), WF_MEM_HOST, None))
    return out


def rough2d_solve_noalloc(x, y, incidences, n_phi: int = 3600, distance: float = 1.0,
                          radius: float = 10e-6, out=None, stream=None):
    """wf_rough2d_solve_noalloc: the device path of ``rough2d_solve`` with no
    device allocation inside the library (its scratch is a fixed static device
    array; see the C header for the size limit).  x, y: float64 CUDA tensors;
    ``out`` (optional) a float64 CUDA tensor [n_inc, n_phi] to write into."""
    import torch

    from ._lib import check, lib

    incidences = list(incidences)
    arr = _incs(incidences)
    k = len(incidences)
    if x.dtype != torch.float64 or y.dtype != torch.float64 or not (x.is_cuda and y.is_cuda):
        raise ValueError("x and y must be float64 CUDA tensors")
    if x.shape != y.shape or x.dim() != 1 or not (x.is_contiguous() and y.is_contiguous()):
        raise ValueError("x and y must be contiguous 1-D tensors of equal length")
    if out is None:
        out = torch.empty((k, n_phi), dtype=torch.float64, device=x.device)
    elif out.shape != (k, n_phi) or out.dtype != torch.float64 or not out.is_contiguous():
        raise ValueError("out must be a contiguous float64 tensor [n_inc, n_phi]")
    st = stream if stream is not None else torch.cuda.current_stream(x.device).cuda_stream
    check(lib().wf_rough2d_solve_noalloc(ct.c_void_p(x.data_ptr()), ct.c_void_p(y.data_ptr()),
                                         x.shape[0], arr, k, n_phi, distance, radius,
                                         ct.c_void_p(out.data_ptr()), ct.c_void_p(st)))
    return out
