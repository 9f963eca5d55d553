"""Reference-compatible command-line drivers.

``python -m wavefiber3d_amd.cli treepo LSTART LEND THETAI PHII FIBER DEPTH START END``
takes the positional arguments of fibersim/treePO.cu:27-34 with the same
meaning -- wavelength indices [LSTART, LEND), incident theta and phi indices,
fiber index into the built-in tables (treePO.cu:45-52), tree depth, and the
arc window [START, END] of each ring -- reads ``mesh/fiber<FIBER>.binary``
(treePO.cu:79-88), converts the heightfield on the GPU, and writes, per
wavelength, ``output/fiber_<FIBER>_lambda<l>_TM_depth<DEPTH>.binary``: the
float32 intensity grid [thetao][phio] of the TM incidence, as
treePO.cu:955-957.  The incidence, wavelength and IOR follow treePO.cu:501-604
(lambda = 400 + 300/25 l nm, thetai = -pi/2 + 1e-10 + pi/180 t, phii =
2 pi/360 p, all rounded to float32 as there; IOR 1.55; Gaussian beam of the
fiber's sigma; D = 1 m) and the octree box is +-1.1 (r1, r2, period/2)
(treePO.cu:94-99).  The tree's own sampling rule and kernels are this
library's (the output grid and file layout are the reference's).

``python -m wavefiber3d_amd.cli sphere LSTART LEND DEPTH`` and
``python -m wavefiber3d_amd.cli ellipsoid DEPTH`` are the drivers of
compareMie/sphere.cu and compareBEM3d/ellipsoid.cu (arguments, mesh files,
incidence, grids and output file names; see ``sphere`` and ``ellipsoid``).
Every tree driver takes ``--reference-sampling``: the reference's level grids
(its beta = 6 and extent x rho rule at 0.4 um) instead of this library's rule.

``python -m wavefiber3d_amd.cli mie`` writes the Mie curves of
compareMie/mie.ipynb cells 2-3 (see ``mie_curves``);
``python -m wavefiber3d_amd.cli rough2d`` is compareBEM2d/rough2d.cu's driver
(see ``rough2d``).

Options of treepo: ``--mesh-dir`` and ``--output-dir`` (defaults ``mesh``, ``output``),
``--final`` also writes ``..._final.binary``, the 180 x 360 block average of
util/fields.h:206-234 (average_cpu) that the reference allocates for each
incidence; ``--fiber-table FILE`` replaces the built-in table with a JSON list
of {period, r1, r2, sigma, numel, numelvert, thetao, phio}; ``--device``.
"""
from __future__ import annotations

import argparse
import json
import os
import sys

import numpy as np

# treePO.cu:45-52
FIBER_TABLE = [
    dict(period=20e-6, r1=20e-6, r2=20e-6, sigma=10e-6, numel=3000, numelvert=478, thetao=450,
         phio=880),
    dict(period=50e-6, r1=40e-6, r2=40e-6, sigma=10e-6, numel=5026, numelvert=1000, thetao=900,
         phio=1440),
    dict(period=400e-6, r1=40e-6, r2=35e-6, sigma=150e-6, numel=4420, numelvert=8000, thetao=7200,
         phio=1440),
    dict(period=400e-6, r1=40e-6, r2=35e-6, sigma=150e-6, numel=4720, numelvert=8000, thetao=7200,
         phio=1440),
]


def f32(x: float) -> float:
    return float(np.float32(x))


def wavelength(index: int) -> float:
    """treePO.cu:504-505: float arithmetic, then *= 1e-9 through double."""
    w = np.float32(np.float32(400.0) + np.float32(300.0) / np.float32(25.0) * np.float32(index))
    return f32(float(w) * 1e-9)


def incidence_angles(theta_index: int, phi_index: int) -> tuple[float, float]:
    """treePO.cu:560, 564 (computed in double, stored float)."""
    import math

    return (f32(-math.pi / 2 + 1e-10 + math.pi / 180 * theta_index),
            f32(2 * math.pi / 360 * phi_index))


def output_name(outdir: str, fiber: int, lam: int, mode: str, depth: int, final: bool = False) -> str:
    """treePO.cu:955."""
    tail = "_final" if final else ""
    return os.path.join(outdir, f"fiber_{fiber}_lambda{lam}_T{mode}_depth{depth}{tail}.binary")


def treepo(argv) -> int:
    ap = argparse.ArgumentParser(prog="python -m wavefiber3d_amd.cli treepo",
                                 description=__doc__.split("\n\n")[0])
    for a in ("lambdaindexstart", "lambdaindexend", "thetaindexstart", "phiiindexstart",
              "fiberindex", "max_depth", "startindex", "endindex"):
        ap.add_argument(a, type=int)
    ap.add_argument("--mesh-dir", default="mesh")
    ap.add_argument("--output-dir", default="output")
    ap.add_argument("--final", action="store_true")
    ap.add_argument("--fiber-table", default=None)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--reference-sampling", action="store_true",
                    help="the reference's level grids: beta 6, rho 1.2 at 0.4 um on the period "
                         "(theta) and 2 max(r1, r2) (phi), treePO.cu:92-93, 342-351")
    a = ap.parse_args(argv)
    table = FIBER_TABLE
    if a.fiber_table:
        with open(a.fiber_table) as fh:
            table = json.load(fh)
    if not 0 <= a.fiberindex < len(table):
        print(f"fiber index {a.fiberindex} outside the table (0..{len(table) - 1})", file=sys.stderr)
        return 2
    p = table[a.fiberindex]

    import torch

    import wavefiber3d_amd as wf

    numel, numelvert = int(p["numel"]), int(p["numelvert"])
    period = f32(p["period"])
    unitz = f32(period / numelvert)
    height = wf.read_heightfield(os.path.join(a.mesh_dir, f"fiber{a.fiberindex}.binary"), numel,
                                 numelvert)
    torch.cuda.set_device(a.device)
    surf = wf.heightfield_samples(torch.from_numpy(height).cuda(), numel, numelvert, period, unitz,
                                  a.startindex, a.endindex)
    # treePO.cu:94-97: float enlargeratio = 1.1; r * enlargeratio in float,
    # period * 0.5 * enlargeratio through double
    enl = np.float32(1.1)
    xr = float(np.float32(p["r1"]) * enl)
    yr = float(np.float32(p["r2"]) * enl)
    zr = f32(period * 0.5 * float(enl))
    os.makedirs(a.output_dir, exist_ok=True)
    thetai, phii = incidence_angles(a.thetaindexstart, a.phiiindexstart)
    grid = (int(p["thetao"]), int(p["phio"]))
    kw = {}
    if a.reference_sampling:
        rho = np.float32(1.2)
        kw = dict(beta=6.0, lambda_sample_m=0.4e-6,
                  ref_sampling=(float(period * rho),
                                float(np.float32(2.0) * np.float32(max(p["r1"], p["r2"])) * rho)))
    with wf.POSolver(device=a.device, box=(-xr, -yr, -zr, xr, yr, zr), **kw) as solver:
        solver.set_geometry(surf, a.max_depth)
        for lam in range(a.lambdaindexstart, a.lambdaindexend):
            inc = wf.Incidence(thetai, phii, wavelength(lam), "M", 1.55 + 0j, f32(p["sigma"]))
            I = solver.solve([inc], grid, out="torch")[0]
            I.cpu().numpy().astype(np.float32).tofile(
                output_name(a.output_dir, a.fiberindex, lam, "M", a.max_depth))
            if a.final:
                wf.average_grid(I, 180, 360).cpu().numpy().tofile(
                    output_name(a.output_dir, a.fiberindex, lam, "M", a.max_depth, final=True))
            print(f"lambdaindex {lam} done", flush=True)
    return 0


def _read_soa_mesh(mesh_dir: str, stem: str):
    """``<stem>_float.binary`` (positions [x...|y...|z...]), ``<stem>_normal_float.binary``
    and ``<stem>_area_float.binary`` in float32 (compareMie/sphere.cu:92-125,
    compareBEM3d/ellipsoid.cu:84-114); the sample count is taken from the area file."""
    import wavefiber3d_amd as wf

    pts = os.path.join(mesh_dir, f"{stem}_float.binary")
    nrm = os.path.join(mesh_dir, f"{stem}_normal_float.binary")
    area = os.path.join(mesh_dir, f"{stem}_area_float.binary")
    for f in (pts, nrm, area):
        if not os.path.exists(f):
            raise FileNotFoundError(f)
    return wf.read_soa_mesh(pts, nrm, area)


def _solve_both_pols(surf, box, depth, grid, lam, eta, sigma, outname, device, ref_sampling):
    """The reference's polarization loop (TM then TE) at incidence theta = phi = 0
    (along +x), one output file per polarization."""
    import wavefiber3d_amd as wf

    kw = {}
    if ref_sampling is not None:
        kw = dict(beta=6.0, lambda_sample_m=0.4e-6, ref_sampling=ref_sampling)
    with wf.POSolver(device=device, box=box, **kw) as solver:
        solver.set_geometry(surf, depth)
        incs = [wf.Incidence(0.0, 0.0, lam, m, eta, sigma) for m in "ME"]
        I = solver.solve(incs, grid)
    for k, m in enumerate("ME"):
        I[k].astype(np.float32).tofile(outname(m))
    return I


def sphere(argv) -> int:
    """compareMie/sphere.cu:21-966: ``sphere LSTART LEND DEPTH`` reads the 10 um
    sphere mesh (``<mesh-dir>/sphere10um_float.binary``, ``..._normal_float``,
    ``..._area_float``; the reference reads 401,182 samples) and writes, per
    wavelength index l in [LSTART, LEND) (lambda = 400 + 300/25 l nm in float)
    and polarization, ``<output-dir>/sphere_10um_lambda<l>_T<M|E>_depth<DEPTH>.binary``:
    the float32 intensity [900][1440] for incidence along +x (theta = phi = 0),
    IOR 1.55, Gaussian beam sigma 1 cm, octree box +-1.1 r (sphere.cu:36-66, 544-590,
    929).  ``--reference-sampling`` takes the reference's level grids (beta 6,
    rho 1.3 on the diameter at 0.4 um, sphere.cu:345-360) instead of the library's rule."""
    ap = argparse.ArgumentParser(prog="python -m wavefiber3d_amd.cli sphere")
    for a_ in ("lambdaindexstart", "lambdaindexend", "max_depth"):
        ap.add_argument(a_, type=int)
    ap.add_argument("--mesh-dir", default="mesh")
    ap.add_argument("--output-dir", default="output")
    ap.add_argument("--radius-um", type=int, default=10)
    ap.add_argument("--reference-sampling", action="store_true")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    try:
        surf = _read_soa_mesh(a.mesh_dir, f"sphere{a.radius_um}um")
    except FileNotFoundError as e:
        print(f"missing mesh file {e}", file=sys.stderr)
        return 2
    radius = f32(a.radius_um * 1e-6)
    r = float(np.float32(radius) * np.float32(1.1))
    box = (-r, -r, -r, r, r, r)
    ref = None
    if a.reference_sampling:
        d = float(np.float32(radius * 2) * np.float32(1.3))
        ref = (d, d)
    os.makedirs(a.output_dir, exist_ok=True)
    for lam in range(a.lambdaindexstart, a.lambdaindexend):
        _solve_both_pols(
            surf, box, a.max_depth, (900, 1440), wavelength(lam), 1.55 + 0j, f32(10000e-6),
            lambda m: os.path.join(a.output_dir,
                                   f"sphere_{a.radius_um}um_lambda{lam}_T{m}_depth{a.max_depth}.binary"),
            a.device, ref)
        print(f"lambdaindex {lam} done", flush=True)
    return 0


def ellipsoid(argv) -> int:
    """compareBEM3d/ellipsoid.cu:20-990: ``ellipsoid DEPTH`` reads the
    3 x 3 x 2 um ellipsoid mesh (``<mesh-dir>/PO/ellipsoid_3_3_2_float.binary``,
    ``..._normal_float``, ``..._area_float``: 30,322 samples) and writes
    ``<output-dir>/ellipsoid_lambda0_T<M|E>_depth<DEPTH>_kappa0.3.binary``: the
    float32 intensity [900][1440] at 500 nm for incidence along +x, IOR
    1.55 + 0.3i, Gaussian beam sigma 1 cm, octree box +-(10, 10, 5) um
    (ellipsoid.cu:38-56, 486-590, 956).  ``--reference-sampling``: the
    reference's level grids (beta 6, rho 3 at 0.4 um, ellipsoid.cu:342-352)."""
    ap = argparse.ArgumentParser(prog="python -m wavefiber3d_amd.cli ellipsoid")
    ap.add_argument("max_depth", type=int)
    ap.add_argument("--mesh-dir", default="mesh")
    ap.add_argument("--output-dir", default="output")
    ap.add_argument("--reference-sampling", action="store_true")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    try:
        surf = _read_soa_mesh(os.path.join(a.mesh_dir, "PO"), "ellipsoid_3_3_2")
    except FileNotFoundError as e:
        print(f"missing mesh file {e}", file=sys.stderr)
        return 2
    r1 = r2 = f32(10e-6)
    period = f32(10e-6)
    zr = f32(period * 0.5)
    box = (-r1, -r2, -zr, r1, r2, zr)
    ref = None
    if a.reference_sampling:
        ref = (float(np.float32(period) * np.float32(3.0)),
               float(np.float32(2.0) * np.float32(max(r1, r2)) * np.float32(3.0)))
    os.makedirs(a.output_dir, exist_ok=True)
    _solve_both_pols(
        surf, box, a.max_depth, (900, 1440), f32(500e-9), 1.55 + 0.3j, f32(10000e-6),
        lambda m: os.path.join(a.output_dir,
                               f"ellipsoid_lambda0_T{m}_depth{a.max_depth}_kappa0.3.binary"),
        a.device, ref)
    print("lambdaindex 0 done", flush=True)
    return 0


def mie_curves(argv) -> int:
    """compareMie/mie.ipynb cells 2-3: the Mie scattering function of a sphere on
    the -180..180 degree grid, written as float64 to
    ``<output-dir>/mie_<radius/1000>um_{TM,TE,TEM}.bin`` (|S1|^2, |S2|^2 and their
    mean; normalisation "t")."""
    ap = argparse.ArgumentParser(prog="python -m wavefiber3d_amd.cli mie")
    ap.add_argument("--wavelength", type=float, default=400.0, help="nm")
    ap.add_argument("--radius", type=float, default=10000.0, help="nm")
    ap.add_argument("--ior", type=complex, default=1.55 + 0.1j)
    ap.add_argument("--resolution", type=float, default=0.25, help="degrees")
    ap.add_argument("--output-dir", default="output")
    a = ap.parse_args(argv)
    from wavefiber3d_amd import mie

    _, SL, SR, SU = mie.scattering_function(a.ior, a.wavelength, 2.0 * a.radius, min_angle=-180,
                                            max_angle=180, angular_resolution=a.resolution,
                                            normalization="t")
    os.makedirs(a.output_dir, exist_ok=True)
    tag = f"{a.radius / 1000:g}um"
    for name, v in (("TM", SL), ("TE", SR), ("TEM", SU)):
        v.astype(np.float64).tofile(os.path.join(a.output_dir, f"mie_{tag}_{name}.bin"))
    print(f"wrote mie_{tag}_{{TM,TE,TEM}}.bin ({SL.shape[0]} angles)", flush=True)
    return 0


def rough2d(argv) -> int:
    """compareBEM2d/rough2d.cu:393-583: ``rough2d TESTCASE MODE LSTART LEND``
    (TESTCASE 0: circle, 1000 samples; otherwise the ellipse, 2000 samples;
    MODE M or E; wavelength indices [LSTART, LEND) of 400 + 300/25 l nm) reads
    ``<crosssection-dir>/circle_{x,y}.txt`` or ``ellipse_new_{x,y}.txt`` (falling
    back to ``ellipse_{x,y}.txt``, the name the reference's data ships under),
    and writes ``<output-dir>/<geo>_T<MODE>_<l>.binary``: 3600 float64 values,
    the scattering at phi = 2 pi k / 3600 for incidence along -x (theta 0,
    phi pi), IOR 1.55 + 0.1i, D = 1 m, normalised by 2 x 10 um."""
    ap = argparse.ArgumentParser(prog="python -m wavefiber3d_amd.cli rough2d")
    ap.add_argument("testcase", type=int)
    ap.add_argument("mode", choices=["M", "E"])
    ap.add_argument("lambdaindexstart", type=int)
    ap.add_argument("lambdaindexend", type=int)
    ap.add_argument("--crosssection-dir", default="crosssection")
    ap.add_argument("--output-dir", default="output")
    a = ap.parse_args(argv)
    from wavefiber3d_amd import rough2d as r2

    if a.testcase == 0:
        geo, names, numel = "circle", ["circle"], 1000
    else:
        geo, names, numel = "ellipse", ["ellipse_new", "ellipse"], 2000
    for nm in names:
        xf = os.path.join(a.crosssection_dir, f"{nm}_x.txt")
        yf = os.path.join(a.crosssection_dir, f"{nm}_y.txt")
        if os.path.exists(xf) and os.path.exists(yf):
            break
    else:
        print(f"no cross-section files for {geo} in {a.crosssection_dir}", file=sys.stderr)
        return 2
    x, y = r2.read_crosssection(xf, yf)
    if x.shape[0] < numel:
        print(f"{xf}: {x.shape[0]} values, the reference reads {numel}", file=sys.stderr)
        return 2
    x, y = x[:numel], y[:numel]
    lams = list(range(a.lambdaindexstart, a.lambdaindexend))
    if not lams:
        return 0
    incs = [r2.Rough2dIncidence(r2.wavelength(l), a.mode) for l in lams]
    I = r2.rough2d_solve(x, y, incs)
    os.makedirs(a.output_dir, exist_ok=True)
    for i, l in enumerate(lams):
        I[i].astype(np.float64).tofile(os.path.join(a.output_dir, f"{geo}_T{a.mode}_{l}.binary"))
        print(f"wavelength index {l} done", flush=True)
    return 0


COMMANDS = {"treepo": treepo, "sphere": sphere, "ellipsoid": ellipsoid, "mie": mie_curves,
            "rough2d": rough2d}


def main(argv=None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    if not argv or argv[0] not in COMMANDS:
        print("usage: python -m wavefiber3d_amd.cli treepo LSTART LEND THETAI PHII FIBER DEPTH "
              "START END [--mesh-dir D] [--output-dir D] [--final] [--fiber-table F] [--device N]\n"
              "       python -m wavefiber3d_amd.cli sphere LSTART LEND DEPTH [--mesh-dir D] "
              "[--output-dir D] [--reference-sampling] [--device N]\n"
              "       python -m wavefiber3d_amd.cli ellipsoid DEPTH [--mesh-dir D] [--output-dir D] "
              "[--reference-sampling] [--device N]\n"
              "       python -m wavefiber3d_amd.cli mie [--wavelength NM] [--radius NM] [--ior C] "
              "[--resolution DEG] [--output-dir D]\n"
              "       python -m wavefiber3d_amd.cli rough2d TESTCASE MODE LSTART LEND "
              "[--crosssection-dir D] [--output-dir D]", file=sys.stderr)
        return 2
    return COMMANDS[argv[0]](argv[1:])


if __name__ == "__main__":
    sys.exit(main())
