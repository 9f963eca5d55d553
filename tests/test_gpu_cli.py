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
