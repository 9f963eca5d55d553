"""Reference-compatible driver on the GPU: a synthetic heightfield file in a
mesh directory, a small fiber table, two wavelengths; the written files have
the reference's names and sizes, hold exactly what the library API returns for
the same inputs, the per-incidence 180 x 360 average is bit-identical to the
block-average oracle, and the intensities match the FP64 direct sum."""
import json
import os

import numpy as np
import pytest

import wavefiber3d_amd as wf
from oracle import oracle
from wavefiber3d_amd import cli

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_treepo_driver(tmp_path):
    r1, r2, length, spacing = 3e-6, 2.5e-6, 2e-6, 0.1e-6
    hx, hy, nv = wf.geometry.fiber_heightfield(r1, r2, length, spacing, seed=5)
    numel = hx.shape[0]
    mesh = os.path.join(tmp_path, "mesh")
    out = os.path.join(tmp_path, "output")
    os.makedirs(mesh)
    np.concatenate([hx.astype(np.float32).ravel(), hy.astype(np.float32).ravel()]).tofile(
        os.path.join(mesh, "fiber0.binary"))
    table = [dict(period=length, r1=r1, r2=r2, sigma=10e-6, numel=numel, numelvert=nv,
                  thetao=360, phio=720)]
    tpath = os.path.join(tmp_path, "table.json")
    with open(tpath, "w") as fh:
        json.dump(table, fh)
    start, end = 10, numel - 20
    print(f"case: treepo driver numel={numel} numelvert={nv}", flush=True)
    rc = cli.main(["treepo", "3", "5", "90", "0", "0", "2", str(start), str(end), "--mesh-dir", mesh,
                   "--output-dir", out, "--final", "--fiber-table", tpath])
    assert rc == 0
    # the same computation through the API
    h = np.fromfile(os.path.join(mesh, "fiber0.binary"), dtype=np.float32)
    period = cli.f32(length)
    s = wf.heightfield_samples(h, numel, nv, period, cli.f32(period / nv), start, end)
    enl = np.float32(1.1)
    xr, yr = float(np.float32(r1) * enl), float(np.float32(r2) * enl)
    zr = cli.f32(period * 0.5 * float(enl))
    thetai, phii = cli.incidence_angles(90, 0)
    with wf.POSolver(device=0, box=(-xr, -yr, -zr, xr, yr, zr)) as sv:
        sv.set_geometry(s, 2)
        for lam in (3, 4):
            f = cli.output_name(out, 0, lam, "M", 2)
            got = np.fromfile(f, dtype=np.float32).reshape(360, 720)
            inc = wf.Incidence(thetai, phii, cli.wavelength(lam), "M", 1.55 + 0j, cli.f32(10e-6))
            I = sv.solve([inc], (360, 720))[0]
            assert np.array_equal(got, I)
            fin = np.fromfile(cli.output_name(out, 0, lam, "M", 2, final=True), dtype=np.float32)
            assert np.array_equal(fin.reshape(180, 360), oracle.average_grid(got, 180, 360))
            Iref, _ = oracle.solve_direct(s.arrays(), [inc.as_tuple()], 360, 720,
                                          centre=(0.0, 0.0, 0.0))
            rel = np.linalg.norm(got.ravel() - Iref[0].ravel()) / np.linalg.norm(Iref[0])
            assert rel < 1e-3, rel


def test_average_grid_device_and_host_bit_identical():
    g = torch.rand((3, 720, 1443), dtype=torch.float32, device="cuda")
    print("case: average grid", flush=True)
    dev = wf.average_grid(g, 180, 360).cpu().numpy()
    host = wf.average_grid(g.cpu().numpy(), 180, 360)
    for k in range(3):
        ref = oracle.average_grid(g[k].cpu().numpy(), 180, 360)
        assert np.array_equal(dev[k], ref)
        assert np.array_equal(host[k], ref)


def test_rough2d_command_writes_reference_files(tmp_path):
    import os
    import shutil

    from oracle import oracle
    from wavefiber3d_amd import cli
    from wavefiber3d_amd import rough2d as r2

    g = os.path.join(os.path.dirname(__file__), "golden")
    xs = os.path.join(tmp_path, "crosssection")
    os.makedirs(xs)
    for f in ("circle_x.txt", "circle_y.txt", "ellipse_x.txt", "ellipse_y.txt"):
        shutil.copy(os.path.join(g, f), xs)
    out = os.path.join(tmp_path, "output")
    print("case: cli rough2d circle TM 3..5, ellipse TE 0..1", flush=True)
    assert cli.main(["rough2d", "0", "M", "3", "5", "--crosssection-dir", xs, "--output-dir", out]) == 0
    assert cli.main(["rough2d", "1", "E", "0", "1", "--crosssection-dir", xs, "--output-dir", out]) == 0
    x, y = r2.read_crosssection(os.path.join(xs, "circle_x.txt"), os.path.join(xs, "circle_y.txt"))
    for l in (3, 4):
        got = np.fromfile(os.path.join(out, f"circle_TM_{l}.binary"), np.float64)
        ref = oracle.rough2d(x, y, 0.0, np.pi, r2.wavelength(l), "M", 1.55 + 0.1j)
        assert got.shape == (3600,)
        assert np.abs(got - ref).max() < 1e-11 * ref.max()
    assert os.path.exists(os.path.join(out, "ellipse_TE_0.binary"))


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def _oracle_dirs(s, incs, I_full, n_dirs=256, seed=0):
    """Tree output vs the C FP64 oracle on n_dirs seeded directions of 900 x 1440."""
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(900 * 1440, n_dirs, replace=False))
    dirs = oracle.output_dirs(900, 1440)[idx]
    Io, _ = oracle.solve_direct(s.arrays(), [t.as_tuple() for t in incs], 0, 0, dirs=dirs)
    return max(_rel(I_full[k].ravel()[idx], Io[k]) for k in range(len(incs)))


def test_ellipsoid_driver_on_reference_mesh(tmp_path):
    """compareBEM3d/ellipsoid.cu's driver on the reference's own 30,322-sample
    mesh: file names and sizes, TM and TE within 1e-4 of the FP64 direct sum
    on the full 900 x 1440 grid and of the C oracle on 256 directions."""
    import shutil

    g = os.path.join(os.path.dirname(__file__), "golden")
    po = os.path.join(tmp_path, "mesh", "PO")
    os.makedirs(po)
    for f in ("ellipsoid_3_3_2_float.binary", "ellipsoid_3_3_2_normal_float.binary",
              "ellipsoid_3_3_2_area_float.binary"):
        shutil.copy(os.path.join(g, f), po)
    out = os.path.join(tmp_path, "output")
    print("case: cli ellipsoid depth 3", flush=True)
    assert cli.main(["ellipsoid", "3", "--mesh-dir", os.path.join(tmp_path, "mesh"),
                     "--output-dir", out]) == 0
    s = cli._read_soa_mesh(po, "ellipsoid_3_3_2")
    assert s.n == 30322
    incs = [wf.Incidence(0.0, 0.0, cli.f32(500e-9), m, 1.55 + 0.3j, cli.f32(10000e-6)) for m in "ME"]
    got = [np.fromfile(os.path.join(out, f"ellipsoid_lambda0_T{m}_depth3_kappa0.3.binary"),
                       np.float32).reshape(900, 1440) for m in "ME"]
    with wf.POSolver(device=0) as sv:
        I64 = sv.solve_fp64_direct(s, incs, (900, 1440))
    for k in range(2):
        assert _rel(got[k], I64[k]) < 1e-4
    assert _oracle_dirs(s, incs, got) < 1e-4


def test_sphere_driver_file_names_and_parity(tmp_path):
    """compareMie/sphere.cu's driver: a 401,182-sample Fibonacci sphere of
    radius 10 um written in the reference's mesh layout (the reference's own
    point and normal files are not shipped), two wavelengths x TM/TE at depth 3;
    file names and sizes, and each grid within 1e-4 of the FP64 direct sum
    and of the C oracle on 256 directions."""
    s = wf.fibonacci_sphere(401_182, 10e-6)
    mesh = os.path.join(tmp_path, "mesh")
    os.makedirs(mesh)
    x, y, z, nx, ny, nz, a = s.arrays()
    np.concatenate([x, y, z]).astype(np.float32).tofile(os.path.join(mesh, "sphere10um_float.binary"))
    np.concatenate([nx, ny, nz]).astype(np.float32).tofile(
        os.path.join(mesh, "sphere10um_normal_float.binary"))
    a.astype(np.float32).tofile(os.path.join(mesh, "sphere10um_area_float.binary"))
    out = os.path.join(tmp_path, "output")
    print("case: cli sphere 0 2 3", flush=True)
    assert cli.main(["sphere", "0", "2", "3", "--mesh-dir", mesh, "--output-dir", out]) == 0
    with wf.POSolver(device=0) as sv:
        for lam in (0, 1):
            incs = [wf.Incidence(0.0, 0.0, cli.wavelength(lam), m, 1.55 + 0j, cli.f32(10000e-6))
                    for m in "ME"]
            got = [np.fromfile(os.path.join(out, f"sphere_10um_lambda{lam}_T{m}_depth3.binary"),
                               np.float32).reshape(900, 1440) for m in "ME"]
            I64 = sv.solve_fp64_direct(s, incs, (900, 1440))
            for k in range(2):
                assert _rel(got[k], I64[k]) < 1e-4
            if lam == 0:
                assert _oracle_dirs(s, incs, got) < 1e-4


def test_reference_sampling_rule_level_grids():
    """--reference-sampling: the level grids are the reference's counts
    (treePO.cu:342-351, beta 6, rho 1.2 at 0.4 um) rounded up to 5-smooth
    lengths (phi even).  The reference's theta count covers the z extent
    only, so its grids under-sample wide boxes near the poles; the error
    against the FP64 sum is reported, not gated."""
    s = wf.fiber_segment(8e-6, 6e-6, 4e-6, 0.1e-6, seed=1)
    period, r = np.float32(4e-6), np.float32(8e-6)
    dt = float(period * np.float32(1.2))
    dp = float(np.float32(2.0) * r * np.float32(1.2))
    inc = [wf.Incidence(0.2, 0.4, 0.5e-6, "M", 1.55 + 0.02j)]
    print("case: reference sampling rule, fiber 8x6x4 um depth 3", flush=True)
    with wf.POSolver(device=0, beta=6.0, lambda_sample_m=0.4e-6, ref_sampling=(dt, dp)) as sv:
        sv.set_geometry(s, 3)
        I = sv.solve(inc, (90, 180))
        lv = sv.stats()["levels"]

    def smooth(n, even=False):
        def ok(m):
            for q in (2, 3, 5):
                while m % q == 0:
                    m //= q
            return m == 1
        while not ok(n) or (even and n % 2):
            n += 1
        return n

    for l in range(1, 4):
        nt = smooth(max(4, oracle.ref_level_count(l, dt, 6.0)))
        np_ = smooth(max(8, oracle.ref_level_count(l, dp, 6.0)), even=True)
        assert tuple(lv[l][:2]) == (nt, np_), (l, lv[l], nt, np_)
    Iref, _ = oracle.solve_direct(s.arrays(), [inc[0].as_tuple()], 90, 180)
    assert np.isfinite(I).all()
    print(f"  levels {lv}; rel-L2 vs FP64 oracle {_rel(I[0], Iref[0]):.2e}", flush=True)
