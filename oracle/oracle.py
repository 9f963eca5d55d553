"""ctypes binding of the FP64 C oracle (po_oracle.c) plus numpy helpers.

TEST INFRASTRUCTURE ONLY.  Each wrapper names the reference file:line the C
function restates.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


class Incidence(ct.Structure):
    _fields_ = [
        ("theta", ct.c_double),
        ("phi", ct.c_double),
        ("lambda_m", ct.c_double),
        ("pol", ct.c_char),
        ("eta_re", ct.c_double),
        ("eta_im", ct.c_double),
        ("beam_sigma_m", ct.c_double),
    ]


def _cpu_has_v3() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
        return all(t in flags for t in (" avx2", " fma", " bmi2"))
    except OSError:
        return False


def build() -> None:
    """Compile the oracle with its Makefile (gcc; called by __graft_entry__.build)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    names = (["libpo_oracle_v3.so"] if _cpu_has_v3() else []) + ["libpo_oracle.so"]
    path = None
    for n in names:
        p = os.path.join(_HERE, n)
        if os.path.exists(p):
            path = p
            break
    if path is None:
        raise RuntimeError("oracle library not built; run oracle.build() or make -C oracle")
    L = ct.CDLL(path)
    dp = ct.POINTER(ct.c_double)
    fp = ct.POINTER(ct.c_float)
    L.orc_z0.restype = ct.c_double
    L.orc_fresnel.argtypes = [ct.c_double, ct.c_double, ct.c_double, dp, dp]
    L.orc_incidence_basis.argtypes = [ct.c_double, ct.c_double, ct.c_char, dp, dp, dp]
    L.orc_radiation_factors.argtypes = [ct.c_double, dp, dp]
    L.orc_currents.argtypes = [ct.c_int64] + [fp] * 7 + [ct.POINTER(Incidence), dp]
    L.orc_output_dirs.argtypes = [ct.c_int, ct.c_int, dp]
    L.orc_direct_sum.argtypes = [ct.c_int64, fp, fp, fp, dp, ct.c_double, ct.c_int, dp,
                                 ct.c_int64, dp, dp, ct.c_int]
    L.orc_intensity.argtypes = [ct.c_int64, dp, dp, dp, dp]
    L.orc_solve_direct.argtypes = [ct.c_int64] + [fp] * 7 + [ct.POINTER(Incidence), ct.c_int,
                                                             ct.c_int, ct.c_int, dp, dp, dp, ct.c_int]
    L.orc_solve_direct.restype = ct.c_int
    L.orc_solve_direct_dirs.argtypes = [ct.c_int64] + [fp] * 7 + [ct.POINTER(Incidence), ct.c_int,
                                                                  ct.c_int64, dp, dp, dp, dp, ct.c_int]
    L.orc_solve_direct_dirs.restype = ct.c_int
    L.orc_direct_sum_finite.argtypes = [ct.c_int64, fp, fp, fp, ct.c_double, ct.c_int, dp,
                                        ct.c_int64, dp, ct.c_double, dp, ct.c_int]
    L.orc_solve_finite_dirs.argtypes = [ct.c_int64] + [fp] * 7 + [ct.POINTER(Incidence), ct.c_int,
                                                                  ct.c_int64, dp, ct.c_double, dp,
                                                                  dp, ct.c_int]
    L.orc_solve_finite_dirs.restype = ct.c_int
    L.orc_rough2d.argtypes = [ct.c_int, dp, dp, ct.c_double, ct.c_double, ct.c_double, ct.c_char,
                              ct.c_double, ct.c_double, ct.c_int, ct.c_double, ct.c_double, dp, dp]
    L.orc_octree.argtypes = [ct.c_int64, fp, fp, fp, fp, ct.c_int, ct.POINTER(ct.c_int64),
                             ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int32), fp]
    L.orc_octree.restype = ct.c_int
    L.orc_ref_level_count.argtypes = [ct.c_int, ct.c_float, ct.c_float]
    L.orc_ref_level_count.restype = ct.c_int
    L.orc_trig_interp.argtypes = [ct.c_int, dp, ct.c_double, ct.c_int, ct.c_double, dp]
    L.orc_sphere_interp.argtypes = [ct.c_int, ct.c_int, ct.c_double, dp, ct.c_int, ct.c_int,
                                    ct.c_double, dp]
    _lib = L
    return L


def _d(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


def _f(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_float))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def z0() -> float:
    return lib().orc_z0()


def fresnel(theta_loc: float, eta: complex):
    """util/util.h:365-382 with n0 = 1 -> (rs, rp) complex."""
    rs = np.zeros(2)
    rp = np.zeros(2)
    lib().orc_fresnel(theta_loc, eta.real, eta.imag, _d(rs), _d(rp))
    return complex(rs[0], rs[1]), complex(rp[0], rp[1])


def incidence_basis(theta: float, phi: float, pol: str):
    """fibersim/treePO.cu:580-601 -> (w, E, H)."""
    w, E, H = np.zeros(3), np.zeros(3), np.zeros(3)
    lib().orc_incidence_basis(theta, phi, pol.encode(), _d(w), _d(E), _d(H))
    return w, E, H


def radiation_factors(lambda_m: float):
    """fibersim/treePO.cu:504-509 -> (E_f, H_f) complex."""
    ef, hf = np.zeros(2), np.zeros(2)
    lib().orc_radiation_factors(lambda_m, _d(ef), _d(hf))
    return complex(ef[0], ef[1]), complex(hf[0], hf[1])


def make_incidences(incs):
    """incs: iterable of dicts / tuples (theta, phi, lambda_m, pol, eta, sigma)."""
    arr = (Incidence * len(incs))()
    for i, t in enumerate(incs):
        if isinstance(t, dict):
            t = (t["theta"], t["phi"], t["lambda_m"], t["pol"], t["eta"], t.get("beam_sigma_m", 0.0))
        th, ph, lam, pol, eta, sig = t
        arr[i] = Incidence(th, ph, lam, pol.encode(), complex(eta).real, complex(eta).imag, sig)
    return arr


def currents(surf, inc) -> np.ndarray:
    """util/tangentplane.h:18-80 + fields.h:106-150 -> complex128 [n, 6] (Jx,Jy,Jz,Mx,My,Mz)."""
    x, y, z, nx, ny, nz, a = [_f32(v) for v in surf]
    n = x.shape[0]
    out = np.zeros((n, 12))
    incs = make_incidences([inc])
    lib().orc_currents(n, _f(x), _f(y), _f(z), _f(nx), _f(ny), _f(nz), _f(a), incs, _d(out))
    return out.view(np.complex128).reshape(n, 6)


def output_dirs(ntheta: int, nphi: int) -> np.ndarray:
    """fibersim/treePO.cu:378-391 in FP64 -> [ntheta*nphi, 3]."""
    d = np.zeros((ntheta * nphi, 3))
    lib().orc_output_dirs(ntheta, nphi, _d(d))
    return d


def direct_sum(xyz, C, k: float, dirs, centre=(0.0, 0.0, 0.0), nthreads: int = 0):
    """fields.h:8-38 + treePO.cu:662-670 (depth 0).  C: complex [S, n, 6] -> F [S, D, 6]."""
    x, y, z = [_f32(v) for v in xyz]
    C = np.ascontiguousarray(C, dtype=np.complex128)
    if C.ndim == 2:
        C = C[None]
    S, n = C.shape[0], C.shape[1]
    dirs = np.ascontiguousarray(dirs, dtype=np.float64)
    F = np.zeros((S, dirs.shape[0], 6), dtype=np.complex128)
    cen = np.asarray(centre, dtype=np.float64)
    lib().orc_direct_sum(n, _f(x), _f(y), _f(z), _d(cen), k, S, _d(C.view(np.float64)),
                         dirs.shape[0], _d(dirs), _d(F.view(np.float64)), nthreads)
    return F


def direct_sum_finite(xyz, C, k: float, dirs, dist: float, nthreads: int = 0):
    """ellipsoid_simple.cu:229-261 at ob = dist * s (about the origin), with the
    far-field factor of the currents (D0 = 1 m) removed.  C: [S, n, 6] -> F [S, D, 6]."""
    x, y, z = [_f32(v) for v in xyz]
    C = np.ascontiguousarray(C, dtype=np.complex128)
    if C.ndim == 2:
        C = C[None]
    S, n = C.shape[0], C.shape[1]
    dirs = np.ascontiguousarray(dirs, dtype=np.float64)
    F = np.zeros((S, dirs.shape[0], 6), dtype=np.complex128)
    lib().orc_direct_sum_finite(n, _f(x), _f(y), _f(z), k, S, _d(C.view(np.float64)),
                                dirs.shape[0], _d(dirs), dist, _d(F.view(np.float64)), nthreads)
    return F


def solve_finite(surf, incs, ntheta: int, nphi: int, dist: float, nthreads: int = 0, dirs=None):
    """The finite-distance validator (ellipsoid_simple.cu:384-589 semantics) on the
    output grid (or ``dirs``): (I [n_inc, D], E [n_inc, 3, D])."""
    x, y, z, nx, ny, nz, a = [_f32(v) for v in surf]
    n = x.shape[0]
    arr = make_incidences(incs)
    dirs = output_dirs(ntheta, nphi) if dirs is None else np.ascontiguousarray(dirs, np.float64)
    D = dirs.shape[0]
    I = np.zeros((len(incs), D))
    E = np.zeros((len(incs), 3, D), dtype=np.complex128)
    rc = lib().orc_solve_finite_dirs(n, _f(x), _f(y), _f(z), _f(nx), _f(ny), _f(nz), _f(a), arr,
                                     len(incs), D, _d(dirs), dist, _d(I), _d(E.view(np.float64)),
                                     nthreads)
    if rc != 0:
        raise MemoryError("oracle solve failed")
    return I, E


def intensity(dirs, F):
    """util/fields.h:172-204 -> (I [D], E'' complex [3, D])."""
    dirs = np.ascontiguousarray(dirs, dtype=np.float64)
    F = np.ascontiguousarray(F, dtype=np.complex128)
    D = dirs.shape[0]
    I = np.zeros(D)
    E = np.zeros((3, D), dtype=np.complex128)
    lib().orc_intensity(D, _d(dirs), _d(F.view(np.float64)), _d(I), _d(E.view(np.float64)))
    return I, E


def solve_direct(surf, incs, ntheta: int, nphi: int, centre=(0.0, 0.0, 0.0), nthreads: int = 0,
                 dirs=None):
    """Depth-0 PO solve in FP64 (sphere.cu:592-658, :909-923 with depth 0).

    Returns (I [n_inc, ntheta*nphi] float64, E [n_inc, 3, ntheta*nphi] complex128).
    If ``dirs`` is given, the grid arguments are ignored and dirs [D, 3] is used."""
    x, y, z, nx, ny, nz, a = [_f32(v) for v in surf]
    n = x.shape[0]
    arr = make_incidences(incs)
    cen = np.asarray(centre, dtype=np.float64)
    if dirs is None:
        D = ntheta * nphi
        I = np.zeros((len(incs), D))
        E = np.zeros((len(incs), 3, D), dtype=np.complex128)
        rc = lib().orc_solve_direct(n, _f(x), _f(y), _f(z), _f(nx), _f(ny), _f(nz), _f(a), arr,
                                    len(incs), ntheta, nphi, _d(cen), _d(I),
                                    _d(E.view(np.float64)), nthreads)
    else:
        dirs = np.ascontiguousarray(dirs, dtype=np.float64)
        D = dirs.shape[0]
        I = np.zeros((len(incs), D))
        E = np.zeros((len(incs), 3, D), dtype=np.complex128)
        rc = lib().orc_solve_direct_dirs(n, _f(x), _f(y), _f(z), _f(nx), _f(ny), _f(nz), _f(a),
                                         arr, len(incs), D, _d(dirs), _d(cen), _d(I),
                                         _d(E.view(np.float64)), nthreads)
    if rc != 0:
        raise MemoryError("oracle solve failed")
    return I, E


def octree(xyz, box, depth: int):
    """util/tree.h:289-715 restated.  Returns (perm, begin, end, node_box)."""
    x, y, z = [_f32(v) for v in xyz]
    n = x.shape[0]
    nn = (8 ** (depth + 1) - 1) // 7
    perm = np.zeros(n, dtype=np.int64)
    b = np.zeros(nn, dtype=np.int32)
    e = np.zeros(nn, dtype=np.int32)
    nb = np.zeros((nn, 6), dtype=np.float32)
    bx = _f32(box)
    rc = lib().orc_octree(n, _f(x), _f(y), _f(z), _f(bx), depth,
                          perm.ctypes.data_as(ct.POINTER(ct.c_int64)),
                          b.ctypes.data_as(ct.POINTER(ct.c_int32)),
                          e.ctypes.data_as(ct.POINTER(ct.c_int32)), _f(nb))
    if rc != 0:
        raise ValueError("oracle octree: bad arguments")
    return perm, b, e, nb


def ref_level_count(level: int, extent: float, beta: float = 6.0) -> int:
    """fibersim/treePO.cu:346-351 sampling rule."""
    return lib().orc_ref_level_count(level, extent, beta)


def trig_interp(vals, off_c: float, np_out: int, off_p: float):
    """util/interpolation.h:11-80 semantics on one line (brute-force DFT)."""
    v = np.ascontiguousarray(vals, dtype=np.complex128)
    out = np.zeros(np_out, dtype=np.complex128)
    lib().orc_trig_interp(v.shape[0], _d(v.view(np.float64)), off_c, np_out, off_p,
                          _d(out.view(np.float64)))
    return out


def sphere_interp(pat, ntp: int, npp: int, off_tp: float):
    """Child pattern [ntc, npc] on the half-sample theta grid -> parent [ntp, npp]."""
    pat = np.ascontiguousarray(pat, dtype=np.complex128)
    ntc, npc = pat.shape
    out = np.zeros((ntp, npp), dtype=np.complex128)
    off_tc = -np.pi / 2 + np.pi / (2 * ntc)
    lib().orc_sphere_interp(ntc, npc, off_tc, _d(pat.view(np.float64)), ntp, npp, off_tp,
                            _d(out.view(np.float64)))
    return out


def heightfield_points(height, numel: int, numelvert: int, period: float, unitz: float,
                       start: int = 0, end: int | None = None):
    """orc_heightfield_points (util/geometry.h:78-178): returns (7 float32
    arrays x, y, z, nx, ny, nz, area, first_bad) with first_bad = -1 or the
    loop-order index i*numel + j of the first sample with a zero vertex."""
    if end is None:
        end = numel - 1
    h = np.ascontiguousarray(np.asarray(height, dtype=np.float32).reshape(-1))
    n = (end - start + 1) * numelvert
    out = [np.zeros(n, dtype=np.float32) for _ in range(7)]
    L = lib()
    fn = L.orc_heightfield_points
    fn.restype = ct.c_long
    fn.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_float, ct.c_float, ct.c_int, ct.c_int] + \
        [ct.c_void_p] * 7
    r = fn(h.ctypes.data_as(ct.c_void_p), numel, numelvert, period, unitz, start, end,
           *[a.ctypes.data_as(ct.c_void_p) for a in out])
    return out, (int(r) - 1 if r else -1)


def average_grid(grid, nto: int, npo: int) -> np.ndarray:
    """orc_average_grid (util/fields.h:206-234) of one [nt][np] float32 grid."""
    g = np.ascontiguousarray(np.asarray(grid, dtype=np.float32))
    nt, npp = g.shape
    out = np.empty((nto, npo), dtype=np.float32)
    fn = lib().orc_average_grid
    fn.restype = None
    fn.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_int, ct.c_int, ct.c_void_p]
    fn(g.ctypes.data_as(ct.c_void_p), nt, npp, nto, npo, out.ctypes.data_as(ct.c_void_p))
    return out


def rough2d(x, y, theta: float, phii: float, lambda_m: float, pol: str, eta: complex,
            nphi: int = 3600, dist: float = 1.0, radius: float = 10e-6, currents: bool = False):
    """compareBEM2d/rough2d.cu:237-585 for one incidence -> scattering [nphi]
    (and the unscaled currents [n, 6] complex when ``currents``)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    n = x.shape[0]
    out = np.zeros(nphi)
    C = np.zeros((n, 6), dtype=np.complex128) if currents else None
    eta = complex(eta)
    lib().orc_rough2d(n, _d(x), _d(y), theta, phii, lambda_m, pol.encode(), eta.real, eta.imag,
                      nphi, dist, radius, _d(out), _d(C.view(np.float64)) if currents else None)
    return (out, C) if currents else out
