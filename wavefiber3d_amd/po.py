"""Host-side mirror of WaveFiber3d's PO solver interface (over the C-ABI).

The reference exposes its solver as driver programs (fibersim/treePO.cu,
compareMie/sphere.cu, compareBEM3d/ellipsoid.cu) calling header functions
of util/*.h.  ``POSolver`` exposes the same steps -- geometry + octree
(tree.h:289), currents (tangentplane.h:84), the tree solve (treePO.cu:501-961)
and the FP64 plain method (depth 0) -- as methods over libwfpo.so.

Arrays may be numpy (host memory; the call returns when results are in host
memory) or torch CUDA tensors (device memory; asynchronous on the current
torch stream).  Every entry point raises ``WfError`` with the library's status
and message on failure; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import WF_MEM_DEVICE, WF_MEM_HOST, check, lib
from .geometry import SurfaceSamples

# Reference parameter grids -------------------------------------------------


def reference_lambda(index: int) -> float:
    """Wavelength of reference index l: 400 + 300/25 l nm (treePO.cu:504)."""
    return (400.0 + 300.0 / 25.0 * index) * 1e-9


def reference_incidence_theta(index: int, count: int = 180) -> float:
    """thetai = -pi/2 + 1e-10 + pi t/180 (treePO.cu:560)."""
    return -math.pi / 2 + 1e-10 + math.pi / count * index


def reference_incidence_phi(index: int, count: int = 360) -> float:
    """phii = 2 pi p/360 (treePO.cu:564)."""
    return 2 * math.pi / count * index


@dataclass
class Incidence:
    theta: float
    phi: float
    lambda_m: float
    pol: str = "M"
    eta: complex = 1.55 + 0j
    beam_sigma_m: float = 0.0

    def to_c(self) -> _lib.Incidence:
        return _lib.Incidence(self.theta, self.phi, self.lambda_m, self.pol.encode(),
                              complex(self.eta).real, complex(self.eta).imag, self.beam_sigma_m)

    def as_tuple(self):
        return (self.theta, self.phi, self.lambda_m, self.pol, complex(self.eta), self.beam_sigma_m)


def _incs(incs) -> tuple:
    incs = list(incs)
    arr = (_lib.Incidence * len(incs))()
    for i, t in enumerate(incs):
        if not isinstance(t, Incidence):
            t = Incidence(*t)
        arr[i] = t.to_c()
    return arr, len(incs)


def average_grid(I, n_theta_out: int = 180, n_phi_out: int = 360, stream=None):
    """Block average of intensity grids (wf_average_grid; util/fields.h:206-234
    average_cpu): I [..., n_theta, n_phi] float32, numpy (host path) or CUDA
    tensor (device path, asynchronous on the current stream) ->
    [..., n_theta_out, n_phi_out] with integer ratios n_theta // n_theta_out,
    n_phi // n_phi_out, bit-identical to the reference's float loop."""
    if hasattr(I, "is_cuda") and I.is_cuda:
        import torch

        x = I.detach().to(torch.float32).contiguous()
        nt, npp = x.shape[-2], x.shape[-1]
        ng = x.numel() // (nt * npp)
        out = torch.empty(x.shape[:-2] + (n_theta_out, n_phi_out), dtype=torch.float32,
                          device=x.device)
        st = stream if stream is not None else torch.cuda.current_stream(x.device).cuda_stream
        check(lib().wf_average_grid(ct.c_void_p(x.data_ptr()), ng, nt, npp, n_theta_out,
                                    n_phi_out, ct.c_void_p(out.data_ptr()), WF_MEM_DEVICE,
                                    ct.c_void_p(st)))
        return out
    x = np.ascontiguousarray(np.asarray(I, dtype=np.float32))
    nt, npp = x.shape[-2], x.shape[-1]
    ng = x.size // (nt * npp)
    out = np.empty(x.shape[:-2] + (n_theta_out, n_phi_out), dtype=np.float32)
    check(lib().wf_average_grid(x.ctypes.data_as(ct.c_void_p), ng, nt, npp, n_theta_out, n_phi_out,
                                out.ctypes.data_as(ct.c_void_p), WF_MEM_HOST, None))
    return out


def output_directions(ntheta: int, nphi: int) -> np.ndarray:
    """Output grid directions (treePO.cu:378-391), float64 [ntheta*nphi, 3]."""
    t = -np.pi / 2 + 1e-10 + np.pi / ntheta * np.arange(ntheta)
    p = 2 * np.pi / nphi * np.arange(nphi)
    T, P = np.meshgrid(t, p, indexing="ij")
    return np.stack([np.cos(T) * np.cos(P), np.cos(T) * np.sin(P), np.sin(T)], -1).reshape(-1, 3)


def _is_torch(a) -> bool:
    return type(a).__module__.startswith("torch")


def _ptr(a) -> int:
    if _is_torch(a):
        return a.data_ptr()
    return a.ctypes.data


def _stream_ptr(arrays) -> int | None:
    for a in arrays:
        if _is_torch(a) and a.is_cuda:
            import torch

            return torch.cuda.current_stream(a.device).cuda_stream
    return None


def _make_opts(device: int = -1, precision: str = "bf16x3", beta: float = 4.0, box=None,
               box_enlarge: float = 1.1, lambda_sample_m: float = 0.0, strict_grid: bool = False,
               workspace_cap_bytes: int = 0, field_chunk: int = 0,
               cancel_flag: ct.c_int32 | None = None, use_graph: bool = False,
               solve_streams: int = 0, ref_sampling=None) -> _lib.Opts:
    o = _lib.Opts()
    lib().wf_default_opts(ct.byref(o))
    o.device = device
    o.precision = {"fp32": _lib.WF_PREC_FP32, "bf16x3": _lib.WF_PREC_BF16X3,
                   "fp32_valu": _lib.WF_PREC_FP32_VALU}[precision]
    o.beta = beta
    o.box_enlarge = box_enlarge
    o.lambda_sample_m = lambda_sample_m
    o.strict_grid = int(strict_grid)
    o.workspace_cap_bytes = workspace_cap_bytes
    o.field_chunk = field_chunk
    o.use_graph = int(use_graph)
    o.solve_streams = int(solve_streams)
    if ref_sampling is not None:  # (d_theta, d_phi) of the reference's sampling rule
        o.ref_dtheta_m, o.ref_dphi_m = float(ref_sampling[0]), float(ref_sampling[1])
    if box is not None:
        o.has_box = 1
        for i in range(6):
            o.box[i] = float(box[i])
    if cancel_flag is not None:
        o.cancel_flag = ct.pointer(cancel_flag)
    return o


def _depth_arg(depth) -> int:
    if isinstance(depth, str):
        if depth != "auto":
            raise ValueError("depth must be an int in [0, 10] or 'auto'")
        return _lib.WF_DEPTH_AUTO
    return int(depth)


def _surface_struct(s):
    arrs = s.arrays() if isinstance(s, SurfaceSamples) else list(s)
    if len(arrs) != 7:
        raise ValueError("surface needs x, y, z, nx, ny, nz, area")
    dev = _is_torch(arrs[0]) and arrs[0].is_cuda
    if dev:
        import torch

        arrs = [a.to(torch.float32).contiguous() for a in arrs]
    else:
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in arrs]
    n = int(arrs[0].shape[0])
    if any(int(a.shape[0]) != n for a in arrs):
        raise ValueError("surface arrays differ in length")
    c = _lib.Surface(*[_ptr(a) for a in arrs], n, WF_MEM_DEVICE if dev else WF_MEM_HOST)
    return c, arrs, dev


class POSolver:
    """Handle over the C-ABI (one per host thread)."""

    def __init__(self, device: int = -1, precision: str = "bf16x3", beta: float = 4.0,
                 box=None, box_enlarge: float = 1.1, lambda_sample_m: float = 0.0,
                 strict_grid: bool = False, workspace_cap_bytes: int = 0,
                 field_chunk: int = 0, cancel_flag: ct.c_int32 | None = None,
                 use_graph: bool = False, solve_streams: int = 0, ref_sampling=None):
        o = _make_opts(device, precision, beta, box, box_enlarge, lambda_sample_m, strict_grid,
                       workspace_cap_bytes, field_chunk, cancel_flag, use_graph, solve_streams,
                       ref_sampling)
        self._cancel = cancel_flag
        self._h = ct.c_void_p()
        check(lib().wf_create(ct.byref(self._h), ct.byref(o)))
        self._keep = []

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().wf_destroy(self._h)
            self._h = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -------------------------------------------------------------- geometry
    def _surface(self, s):
        return _surface_struct(s)

    def set_geometry(self, surface, depth):
        """Upload + octree build (tree.h:289-715; treePO.cu:137-337).  depth: int in
        [0, 10] or "auto" (built to ~log8 n; each wavelength group then uses the
        cost model's depth)."""
        c, arrs, dev = self._surface(surface)
        self._keep = arrs
        check(lib().wf_set_geometry(self._h, ct.byref(c), _depth_arg(depth), _stream_ptr(arrs)))
        n = ct.c_int64()
        d = ct.c_int32()
        check(lib().wf_geometry_info(self._h, ct.byref(n), ct.byref(d), None))
        self.depth = d.value  # built depth
        self.auto_depth = depth == "auto"
        self.n = n.value
        self._dev_geom = dev

    # ------------------------------------------------------------- autotuning
    @property
    def tune(self) -> dict:
        t = _lib.Tune()
        check(lib().wf_get_tune(self._h, ct.byref(t)))
        return {k: getattr(t, k) for k, _ in _lib.Tune._fields_}

    @tune.setter
    def tune(self, d: dict):
        t = _lib.Tune()
        check(lib().wf_get_tune(self._h, ct.byref(t)))
        for k, v in d.items():
            setattr(t, k, v)
        check(lib().wf_set_tune(self._h, ct.byref(t)))

    def autotune(self, incidences, grid) -> int:
        """Fit the depth cost model and leaf-kernel choice on the current geometry
        (timed solves of the first wavelength group); returns the fastest depth."""
        arr, k = _incs(incidences)
        best = ct.c_int32()
        st = None
        if self._dev_geom:
            import torch

            st = torch.cuda.current_stream().cuda_stream
        check(lib().wf_autotune(self._h, arr, k, _lib.Grid(*grid), ct.byref(best), st))
        return best.value

    # ----------------------------------------------------------------- solves
    def solve(self, incidences, grid, return_E: bool = False, out: str = "numpy"):
        """Tree (or depth-0) PO solve on the current geometry.

        Returns intensity [n_inc, Nt, Np] float32 (and E [n_inc, 3, Nt, Np]
        complex64 when return_E).  out = "numpy" (host) or "torch" (device)."""
        arr, k = _incs(incidences)
        nt, np_ = grid
        g = _lib.Grid(nt, np_)
        if out == "torch":
            import torch

            I = torch.empty((k, nt, np_), dtype=torch.float32, device="cuda")
            E = torch.empty((k, 3, nt, np_, 2), dtype=torch.float32, device="cuda") if return_E else None
            st = torch.cuda.current_stream().cuda_stream
            check(lib().wf_solve(self._h, arr, k, g, I.data_ptr(), E.data_ptr() if E is not None else None,
                                 WF_MEM_DEVICE, st))
            if return_E:
                return I, torch.view_as_complex(E)
            return I
        I = np.empty((k, nt, np_), dtype=np.float32)
        E = np.empty((k, 3, nt, np_, 2), dtype=np.float32) if return_E else None
        check(lib().wf_solve(self._h, arr, k, g, I.ctypes.data,
                             E.ctypes.data if E is not None else None, WF_MEM_HOST, None))
        if return_E:
            return I, E.view(np.complex64)[..., 0]
        return I

    def po_solve(self, surface, incidences, grid, depth: int, return_E: bool = False):
        self.set_geometry(surface, depth)
        return self.solve(incidences, grid, return_E=return_E,
                          out="torch" if self._dev_geom else "numpy")

    def solve_fp64_direct(self, surface, incidences, grid, return_E: bool = False):
        """FP64 plain method on the GPU (sphere.cu:592-658 with depth 0)."""
        return self._fp64(surface, incidences, grid, return_E, None)

    def solve_fp64_finite(self, surface, incidences, grid, distance: float = 1.0,
                          return_E: bool = False):
        """FP64 direct sum with the exact Green's function exp(-ik|x-ob|)/|x-ob|
        at observation points ob = distance * s (compareBEM3d/ellipsoid_simple.cu:
        229-261, 384-589; the reference uses 1 m).  Same outputs as
        solve_fp64_direct; for distance -> infinity the two agree."""
        return self._fp64(surface, incidences, grid, return_E, float(distance))

    def _fp64(self, surface, incidences, grid, return_E, distance):
        c, arrs, dev = self._surface(surface)
        arr, k = _incs(incidences)
        nt, np_ = grid
        g = _lib.Grid(nt, np_)
        if distance is None:
            call, dist = lib().wf_po_solve_fp64_direct, ()
        else:
            call, dist = lib().wf_po_solve_fp64_finite, (distance,)
        if dev:
            import torch

            I = torch.empty((k, nt, np_), dtype=torch.float64, device=arrs[0].device)
            E = torch.empty((k, 3, nt, np_, 2), dtype=torch.float64, device=arrs[0].device) if return_E else None
            check(call(self._h, ct.byref(c), arr, k, g, *dist, I.data_ptr(),
                       E.data_ptr() if E is not None else None, _stream_ptr(arrs)))
            self.depth = 0
            if return_E:
                return I, torch.view_as_complex(E)
            return I
        I = np.empty((k, nt, np_), dtype=np.float64)
        E = np.empty((k, 3, nt, np_, 2), dtype=np.float64) if return_E else None
        check(call(self._h, ct.byref(c), arr, k, g, *dist, I.ctypes.data,
                   E.ctypes.data if E is not None else None, None))
        self.depth = 0
        if return_E:
            return I, E.view(np.complex128)[..., 0]
        return I

    # ------------------------------------------------------------ stage APIs
    def currents(self, surface_dev, incidences, fp64: bool = False):
        """computeLocalTangentPlane + area + E_f/H_f (device tensors in, [n_inc, n, 6] complex out)."""
        import torch

        c, arrs, dev = self._surface(surface_dev)
        if not dev:
            raise ValueError("currents() needs CUDA tensors")
        arr, k = _incs(incidences)
        dt = torch.float64 if fp64 else torch.float32
        out = torch.empty((k, c.n, 6, 2), dtype=dt, device=arrs[0].device)
        check(lib().wf_currents(self._h, ct.byref(c), arr, k, int(fp64), out.data_ptr(),
                                _stream_ptr(arrs)))
        return torch.view_as_complex(out)

    def octree_nodes(self):
        """(perm [n], node_begin, node_end [(8^(d+1)-1)/7]) as int32 CUDA tensors."""
        import torch

        nn = (8 ** (self.depth + 1) - 1) // 7
        perm = torch.empty(self.n, dtype=torch.int32, device="cuda")
        nb = torch.empty(nn, dtype=torch.int32, device="cuda")
        ne = torch.empty(nn, dtype=torch.int32, device="cuda")
        check(lib().wf_octree_nodes(self._h, perm.data_ptr(), nb.data_ptr(), ne.data_ptr(),
                                    torch.cuda.current_stream().cuda_stream))
        return perm, nb, ne

    def interp(self, child, ntp: int, npp: int, off_tp: float):
        """copyandinterp/interpFFT on a batch of child patterns (complex64 CUDA tensor [count, ntc, npc])."""
        import torch

        child = child.to(torch.complex64).contiguous()
        count, ntc, npc = child.shape
        out = torch.empty((count, ntp, npp), dtype=torch.complex64, device=child.device)
        check(lib().wf_interp(self._h, count, ntc, npc, child.data_ptr(), ntp, npp, off_tp,
                              out.data_ptr(), torch.cuda.current_stream().cuda_stream))
        return out

    GRAD_MODES = {"tree32": 0, "direct64": 1, "tree64": 2}

    def vjp(self, incidences, grid, grad_I, fp64: bool = False, return_I: bool = False,
            mode: str | None = None):
        """Reverse mode on the current geometry: dL/dI -> (dL/dx [3, n], dL/dn [3, n],
        dL/da [n]) in the original point order.  grad_I: numpy (host) or torch CUDA
        [n_inc, Nt, Np].  mode: "tree64" (default: the tree with FP64 patterns),
        "tree32" (FP32 patterns) or "direct64" (the FP64 direct sum, also
        fp64=True)."""
        code = self._grad_mode(fp64, mode)
        arr, k = _incs(incidences)
        nt, np_ = grid
        g = _lib.Grid(nt, np_)
        n = self.n
        if _is_torch(grad_I) and grad_I.is_cuda:
            import torch

            gI = grad_I.to(torch.float32).contiguous()
            dev = gI.device
            gx = torch.empty((3, n), dtype=torch.float32, device=dev)
            gn = torch.empty((3, n), dtype=torch.float32, device=dev)
            ga = torch.empty((n,), dtype=torch.float32, device=dev)
            I = torch.empty((k, nt, np_), dtype=torch.float32, device=dev) if return_I else None
            check(lib().wf_po_vjp(self._h, arr, k, g, gI.data_ptr(),
                                  I.data_ptr() if I is not None else None, gx.data_ptr(),
                                  gn.data_ptr(), ga.data_ptr(), code, WF_MEM_DEVICE,
                                  torch.cuda.current_stream(dev).cuda_stream))
        else:
            gI = np.ascontiguousarray(grad_I, dtype=np.float32)
            gx = np.empty((3, n), np.float32)
            gn = np.empty((3, n), np.float32)
            ga = np.empty((n,), np.float32)
            I = np.empty((k, nt, np_), np.float32) if return_I else None
            check(lib().wf_po_vjp(self._h, arr, k, g, gI.ctypes.data,
                                  I.ctypes.data if I is not None else None, gx.ctypes.data,
                                  gn.ctypes.data, ga.ctypes.data, code, WF_MEM_HOST, None))
        if return_I:
            return gx, gn, ga, I
        return gx, gn, ga

    def _grad_mode(self, fp64: bool, mode: str | None) -> int:
        if fp64:
            mode = "direct64"
        if mode is None:
            mode = "tree64"
        if mode not in self.GRAD_MODES:
            raise ValueError(f"mode must be one of {sorted(self.GRAD_MODES)}")
        return self.GRAD_MODES[mode]

    def grad_forward(self, incidences, grid, mode: str = "tree64", out: str = "torch"):
        """The forward intensity [n_inc, Nt, Np] in the arithmetic of ``vjp(mode=...)``
        (wf_po_vjp with grad_I = NULL), so that a loss and its gradient see the same
        forward."""
        code = self._grad_mode(False, mode)
        arr, k = _incs(incidences)
        nt, np_ = grid
        g = _lib.Grid(nt, np_)
        if out == "torch":
            import torch

            I = torch.empty((k, nt, np_), dtype=torch.float32, device="cuda")
            check(lib().wf_po_vjp(self._h, arr, k, g, None, I.data_ptr(), None, None, None, code,
                                  WF_MEM_DEVICE, torch.cuda.current_stream().cuda_stream))
            return I
        I = np.empty((k, nt, np_), np.float32)
        check(lib().wf_po_vjp(self._h, arr, k, g, None, I.ctypes.data, None, None, None, code,
                              WF_MEM_HOST, None))
        return I

    def stats(self) -> dict:
        """Per-kernel HIP-event timings of the last solve (waits for its events)."""
        s = _lib.Stats()
        check(lib().wf_get_stats(self._h, ct.byref(s)))
        nl = s.n_levels
        kern = {}
        for i, name in enumerate(_lib.KERNEL_KINDS):
            if s.kern_launches[i]:
                kern[name] = {"ms": s.kern_ms[i], "launches": s.kern_launches[i],
                              "work": s.kern_work[i]}
        return {
            "ms_total": s.ms_total, "kernels": kern, "leaf_pairs": s.leaf_pairs, "depth": s.depth,
            "levels": [(s.level_nt[i], s.level_np[i], s.level_nodes[i]) for i in range(nl)],
            "agg_transitions": s.agg_transitions, "agg_bytes": s.agg_bytes,
        }


class Batch:
    """Many surfaces solved for the same incidences and grid (wf_batch_*).

    Each segment keeps its own octree; solves run on ``n_streams`` worker
    streams with separate workspaces.  ``depth="auto"`` segments get their
    depth per wavelength group from the cost model, calibrated on the largest
    one at the first solve."""

    def __init__(self, n_streams: int = 4, **opts):
        o = _make_opts(**opts)
        self._b = ct.c_void_p()
        check(lib().wf_batch_create(ct.byref(self._b), ct.byref(o), int(n_streams)))
        self.n_points = []

    def close(self):
        if getattr(self, "_b", None) and self._b.value:
            lib().wf_batch_destroy(self._b)
            self._b = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, surface, depth="auto") -> int:
        c, arrs, dev = _surface_struct(surface)
        idx = ct.c_int32()
        if dev:
            import torch

            torch.cuda.current_stream().synchronize()
        check(lib().wf_batch_add(self._b, ct.byref(c), _depth_arg(depth), ct.byref(idx)))
        self.n_points.append(c.n)
        return idx.value

    def __len__(self):
        return len(self.n_points)

    def solve(self, incidences, grid, out: str = "numpy"):
        """Intensity [n_segments, n_inc, Nt, Np] float32."""
        arr, k = _incs(incidences)
        nt, np_ = grid
        g = _lib.Grid(nt, np_)
        shape = (len(self), k, nt, np_)
        if out == "torch":
            import torch

            I = torch.empty(shape, dtype=torch.float32, device="cuda")
            check(lib().wf_batch_solve(self._b, arr, k, g, I.data_ptr(), WF_MEM_DEVICE,
                                       torch.cuda.current_stream().cuda_stream))
            return I
        I = np.empty(shape, dtype=np.float32)
        check(lib().wf_batch_solve(self._b, arr, k, g, I.ctypes.data, WF_MEM_HOST, None))
        return I

    def info(self) -> dict:
        """Per segment: depth of the last wavelength group, solve time (ms), points."""
        n = len(self)
        cnt = ct.c_int32()
        d = np.zeros(n, np.int32)
        ms = np.zeros(n, np.float64)
        npts = np.zeros(n, np.int64)
        check(lib().wf_batch_info(self._b, ct.byref(cnt), d.ctypes.data, ms.ctypes.data,
                                  npts.ctypes.data))
        return {"depth": d, "ms": ms, "n_points": npts}
