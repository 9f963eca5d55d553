"""The benchmark workloads at full size on the GPU, through the C-ABI.

Each tree result is checked two ways:
  * against the HIP FP64 direct sum over the full output grid, every solve
    (relative L2 of intensity <= 1e-4; the tree's observed error is 1e-5 to
    2e-5 at the sampling constant the benchmark uses);
  * against the C FP64 oracle on a seeded subset of output directions
    (the same bar), and the HIP FP64 direct sum against the oracle on that
    subset (<= 1e-12).

Workloads:
  * sphere r = 2 um, N = 12,000, IOR 1.55, +z incidence, TM and TE, 0.5 um,
    90 x 180 (the plain method: FP64 direct sum only);
  * sphere r = 10 um, N = 500,000, IOR 1.55 + 0.1i, depth 4, 25 wavelengths
    400-688 nm x TM/TE, +z incidence, 180 x 360;
  * fiber segment 40 x 30 um, length 30 um (663,000 samples, cuticle 5 um /
    3 deg, noise 0.05 um / 0.5 um seed 0), IOR 1.55 + 0.02i, depth 6,
    32 incidences (theta index 60, 75, 90, 105 x phi index 0..315 step 45),
    TM, 0.4 um, 360 x 720 -- depth-6 tree, 32 incidences in one call, the LDS
    FFT root transition with FP64 accumulation;
  * the uneven multi-segment batch with hipGraph capture on: 8 of the seeded
    fiber-family segments, 8 incidences x 25 wavelengths, 360 x 720; the
    graph replay equals the eager solve bit for bit and every segment's first
    wavelength group is checked against the FP64 direct sum.
"""
import math

import numpy as np
import pytest

import oracle
import wavefiber3d_amd as wf

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

TREE_GATE = 1e-4
BETA = 2.0  # the benchmark's level sampling constant


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _oracle_subset(surf, incs, grid, I_tree, I64, n_dirs=256, max_solves=4, seed=0):
    """Tree and HIP FP64 intensities vs the C oracle on n_dirs seeded output
    directions of the first max_solves solves."""
    nt, np_ = grid
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(nt * np_, n_dirs, replace=False))
    dirs = oracle.output_dirs(nt, np_)[idx]
    ks = list(range(min(max_solves, len(incs))))
    Io, _ = oracle.solve_direct(surf.arrays(), [incs[k].as_tuple() for k in ks], 0, 0, dirs=dirs)
    et = [rel(I_tree[k].reshape(-1)[idx], Io[j]) for j, k in enumerate(ks)]
    e64 = [rel(I64[k].reshape(-1)[idx], Io[j]) for j, k in enumerate(ks)]
    return max(et), max(e64)


def _tree_vs_fp64(surf, incs, grid, depth):
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in surf.arrays()]
    with wf.POSolver(device=0, beta=BETA) as sv:
        sv.set_geometry(dev, depth)
        I = sv.solve(incs, grid, out="torch").cpu().numpy()
        st = sv.stats()
        I64 = sv.solve_fp64_direct(dev, incs, grid).cpu().numpy()
    S = len(incs)
    err = [rel(I[k], I64[k]) for k in range(S)]
    return I, I64, err, st


def test_plain_method_sphere_12k():
    s = wf.fibonacci_sphere(12_000, 2e-6)
    incs = [wf.Incidence(math.pi / 2, 0.0, 0.5e-6, p, 1.55 + 0j) for p in "ME"]
    grid = (90, 180)
    print("case: plain method sphere r=2um n=12000 90x180 TM+TE", flush=True)
    with wf.POSolver(device=0, box=[-2.2e-6] * 3 + [2.2e-6] * 3) as sv:
        I, E = sv.solve_fp64_direct(s, incs, grid, return_E=True)
    Io, Eo = oracle.solve_direct(s.arrays(), [t.as_tuple() for t in incs], *grid)
    for k in range(2):
        assert rel(I[k], Io[k]) < 1e-12
        assert rel(E[k].reshape(3, -1), Eo[k]) < 1e-12


def test_sphere_500k_depth4_25_wavelengths():
    s = wf.fibonacci_sphere(500_000, 10e-6)
    incs = [wf.Incidence(math.pi / 2, 0.0, wf.reference_lambda(l), p, 1.55 + 0.1j)
            for l in range(25) for p in "ME"]
    grid = (180, 360)
    print("case: sphere r=10um n=500000 depth 4, 25 lambda x TM/TE, 180x360", flush=True)
    I, I64, err, st = _tree_vs_fp64(s, incs, grid, 4)
    print(f"  depth {st['depth']} levels {st['levels']}; rel-L2 vs FP64 max {max(err):.2e}",
          flush=True)
    assert st["depth"] == 4
    assert max(err) < TREE_GATE, err
    et, e64 = _oracle_subset(s, incs, grid, I, I64)
    print(f"  on 256 directions vs C oracle: tree {et:.2e}, FP64 {e64:.2e}", flush=True)
    assert et < TREE_GATE
    assert e64 < 1e-12


def test_fiber_663k_depth6_32_incidences():
    s = wf.fiber_segment(40e-6, 30e-6, 30e-6, 0.1e-6, seed=0)
    incs = [wf.Incidence(wf.reference_incidence_theta(t), wf.reference_incidence_phi(p), 0.4e-6,
                         "M", 1.55 + 0.02j) for t in (60, 75, 90, 105) for p in range(0, 360, 45)]
    grid = (360, 720)
    print(f"case: fiber n={s.n} depth 6, 32 incidences, 360x720", flush=True)
    assert abs(s.n - 663_000) < 1_000
    I, I64, err, st = _tree_vs_fp64(s, incs, grid, 6)
    print(f"  depth {st['depth']} levels {st['levels']}; rel-L2 vs FP64 max {max(err):.2e}",
          flush=True)
    assert st["depth"] == 6
    assert max(err) < TREE_GATE, err
    et, e64 = _oracle_subset(s, incs, grid, I, I64, max_solves=2)
    print(f"  on 256 directions vs C oracle: tree {et:.2e}, FP64 {e64:.2e}", flush=True)
    assert et < TREE_GATE
    assert e64 < 1e-12


def test_autotune_picks_beta_under_its_gate_on_the_fiber():
    """At the library defaults (opts.beta = 4) wf_autotune lowers the fiber's
    sampling constant to 2 (its solves stay within 1e-4 of the beta = 4 ones)
    and the tuned solve still meets the FP64 direct sum at 1e-4."""
    s = wf.fiber_segment(40e-6, 30e-6, 30e-6, 0.1e-6, seed=0)
    incs = [wf.Incidence(wf.reference_incidence_theta(t), wf.reference_incidence_phi(p), 0.4e-6,
                         "M", 1.55 + 0.02j) for t in (60, 75, 90, 105) for p in (0, 180)]
    grid = (360, 720)
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in s.arrays()]
    print(f"case: autotune fiber n={s.n} depth 6, 8 incidences, 360x720", flush=True)
    with wf.POSolver(device=0) as sv:
        sv.set_geometry(dev, 6)
        sv.autotune(incs, grid)
        t = sv.tune
        print(f"  tune {t}", flush=True)
        assert t["beta"] == 2.0
        # the FFT threshold only moves down from the default, one level at a time
        assert t["agg_fft_rows"] <= 600 and t["chunk_solves"] in (4, 8)
        I = sv.solve(incs[:2], grid, out="torch").cpu().numpy()
        I64 = sv.solve_fp64_direct(dev, incs[:2], grid).cpu().numpy()
    e = max(rel(I[k], I64[k]) for k in range(2))
    print(f"  rel-L2 vs FP64 {e:.2e}", flush=True)
    assert e < TREE_GATE


def test_fiber_fft_and_mfma_levels_agree():
    """The fiber's 125 x 240 -> 240 x 480 and 72 x 128 -> 125 x 240 transitions
    on the wave-private LDS FFT kernels (agg_fft_rows 479, 249) or on the MFMA
    kernels (600): the same interpolation, so the solves agree far inside the
    tree's error against the FP64 direct sum."""
    s = wf.fiber_segment(40e-6, 30e-6, 30e-6, 0.1e-6, seed=0)
    incs = [wf.Incidence(wf.reference_incidence_theta(t), 0.0, 0.4e-6, "E", 1.55 + 0.02j)
            for t in (60, 90)]
    grid = (360, 720)
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in s.arrays()]
    out = {}
    with wf.POSolver(device=0, beta=BETA) as sv:
        sv.set_geometry(dev, 6)
        for rows in (600, 479, 249):
            print(f"case: fiber depth 6, agg_fft_rows {rows}", flush=True)
            sv.tune = {"agg_fft_rows": rows}
            out[rows] = sv.solve(incs, grid, out="torch").cpu().numpy()
    for rows in (479, 249):
        e = max(rel(out[rows][j], out[600][j]) for j in range(len(incs)))
        print(f"  agg_fft_rows {rows}: max rel-L2 vs the MFMA levels {e:.2e}", flush=True)
        assert e < 3e-5


def test_fiber_two_pass_tiles_agree():
    """The two-pass transition of 6-9 K steps (the fiber's 125 x 240 ->
    240 x 480 level) with each theta block tile and phi row tile gives the
    default's result: the tiles only change which block computes a sample."""
    s = wf.fiber_segment(40e-6, 30e-6, 30e-6, 0.1e-6, seed=0)
    incs = [wf.Incidence(wf.reference_incidence_theta(t), 0.0, 0.4e-6, "M", 1.55 + 0.02j)
            for t in (60, 75, 90, 105)]
    grid = (360, 720)
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in s.arrays()]
    out = {}
    with wf.POSolver(device=0, beta=BETA) as sv:
        sv.set_geometry(dev, 6)
        for tt, pr in ((0, 0), (2, 16), (3, 16), (3, 0)):
            print(f"case: fiber depth 6, theta_tile {tt}, phi_rows {pr}", flush=True)
            sv.tune = {"theta_tile": tt, "phi_rows": pr}
            out[(tt, pr)] = sv.solve(incs, grid, out="torch").cpu().numpy()
    for k, I in out.items():
        e = max(rel(I[j], out[(0, 0)][j]) for j in range(len(incs)))
        print(f"  {k}: max rel-L2 vs default tiles {e:.2e}", flush=True)
        assert e < 1e-6


def test_uneven_batch_8_segments_graph():
    from wavefiber3d_amd.geometry import fiber_family_segment

    incs = [wf.Incidence(wf.reference_incidence_theta(t), wf.reference_incidence_phi(p),
                         wf.reference_lambda(l), "M", 1.55 + 0.02j)
            for l in range(25) for t in (60, 90, 105, 120) for p in (0, 180)]
    grid = (360, 720)
    segs = [fiber_family_segment(k) for k in range(8)]
    print(f"case: batch of 8 fiber segments n={[s.n for s in segs]}, 200 solves each, "
          f"graph on", flush=True)
    with wf.Batch(n_streams=4, device=0, beta=BETA, use_graph=True) as b:
        for s in segs:
            b.add(s, "auto")
        I0 = b.solve(incs, grid, out="numpy")  # eager + capture
        I1 = b.solve(incs, grid, out="numpy")  # replay
        info = b.info()
    print(f"  depths {info['depth'].tolist()}", flush=True)
    assert np.array_equal(I0, I1)
    g0 = incs[:8]  # the first wavelength group
    with wf.POSolver(device=0) as sv:
        for k, s in enumerate(segs):
            I64 = sv.solve_fp64_direct(s, g0, grid)
            e = max(rel(I1[k, j], I64[j]) for j in range(8))
            print(f"  segment {k} n={s.n} depth {info['depth'][k]}: rel-L2 vs FP64 {e:.2e}",
                  flush=True)
            assert e < TREE_GATE


def test_adjoint_fiber_200k_tree_gradient_vs_fd():
    """The adjoint workload: 200,000-point fiber (length 10 um), 180 x 360,
    loss sum (log I - log T)^2.  The tree gradient with FP64 patterns (the
    default gradient mode) against central differences of the FP64 forward
    (64 points along their normals at 1e-4 um and h/2, and the cuticle tilt
    along the surface tangent, Richardson-extrapolated; gate 1e-4) and
    against the FP64 direct-sum gradient."""
    from wavefiber3d_amd import gradcheck as gc

    incs = [wf.Incidence(wf.reference_incidence_theta(90), 0.0, 0.4e-6, "M", 1.55 + 0.02j)]
    grid = (180, 360)
    tau0 = torch.tensor(3.0, dtype=torch.float64, device="cuda")
    s64 = [t.detach() for t in gc.c5_surface(tau0)]
    s32 = [t.float().contiguous() for t in s64]
    nrm = torch.stack(s64[3:6])
    print(f"case: adjoint fiber n={s32[0].numel()} 180x360", flush=True)
    assert s32[0].numel() == 200_000
    with wf.POSolver(device=0) as sv:
        I0 = sv.solve_fp64_direct(s32, incs, grid)
        rng = np.random.default_rng(12345)
        target = (I0 * torch.from_numpy(np.exp(0.3 * rng.standard_normal(I0.shape))).cuda()).float()
        sv.set_geometry(s32, "auto")
        depth = sv.autotune(incs, grid)
        grads = {m: gc.input_gradients(sv, s32, incs, grid, target, m, 0 if m == "direct64" else depth)
                 for m in ("direct64", "tree64")}
        gd = {m: (torch.stack(g[:3]) * nrm).sum(0) for m, g in grads.items()}
        rel = float(torch.linalg.norm(gd["tree64"] - gd["direct64"]) / torch.linalg.norm(gd["direct64"]))
        rn = gc.fd_normal_ratio(sv, s32, nrm, incs, grid, target, grads)
        rt = gc.fd_tangent_ratio(sv, gc.c5_surface, 3.0, incs, grid, target, grads)
    print(f"  depth {depth}: tree64 vs FP64 gradient {rel:.2e}; vs FD normals "
          f"{abs(rn['tree64'] - 1):.2e} (FP64 {abs(rn['direct64'] - 1):.2e}); tilt "
          f"{abs(rt['tree64'] - 1):.2e} (FP64 {abs(rt['direct64'] - 1):.2e})", flush=True)
    assert abs(rn["tree64"] - 1) < 1e-4
    assert abs(rt["tree64"] - 1) < 1e-4
    assert rel < 1e-4
