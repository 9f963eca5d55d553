"""GPU parity tests: each HIP stage through the C-ABI vs the FP64 oracle.

Bars: FP64 paths <= 1e-12 relative; FP32 stages <= 1e-5; tree solves <= 1e-4
relative L2 of intensity against the FP64 oracle (the tree path's gate is
1e-3; its observed error is 2e-6 to 2e-5), plus exact integer parity for the
octree.
"""
import numpy as np
import pytest

import oracle
import wavefiber3d_amd as wf

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(s):
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in s.arrays()]


def _rand_surface(n, seed, scale=2e-6):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-scale, scale, (3, n))
    q = rng.standard_normal((3, n))
    q /= np.linalg.norm(q, axis=0)
    a = rng.uniform(1e-15, 5e-15, n)
    return wf.SurfaceSamples(*[v.astype(np.float32) for v in (p[0], p[1], p[2], q[0], q[1], q[2], a)])


def rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


INCS = [
    wf.Incidence(np.pi / 2, 0.0, 0.5e-6, "M", 1.55 + 0.0j),
    wf.Incidence(np.pi / 2, 0.0, 0.5e-6, "E", 1.55 + 0.0j),
    wf.Incidence(-0.4, 2.1, 0.61e-6, "M", 1.55 + 0.1j, 3e-6),
    wf.Incidence(0.7, -1.3, 0.43e-6, "E", 1.3 + 0.02j),
]


def test_currents_fp64_and_fp32(solver):
    s = _rand_surface(3000, 1)
    print("case: currents n=3000", flush=True)
    C64 = solver.currents(_dev(s), INCS, fp64=True).cpu().numpy()
    C32 = solver.currents(_dev(s), INCS, fp64=False).cpu().numpy()
    for k, inc in enumerate(INCS):
        ref = oracle.currents(s.arrays(), inc.as_tuple())
        assert rel(C64[k], ref) < 1e-12
        assert rel(C32[k], ref) < 1e-6


def test_direct_fp64_matches_oracle():
    s = wf.fibonacci_sphere(3000, 2e-6)
    nt, np_ = 24, 48
    print("case: fp64 direct sphere n=3000 24x48", flush=True)
    # complex E depends on the phase centre (the root box centre): use the oracle's origin
    with wf.POSolver(device=0, box=[-2.2e-6] * 3 + [2.2e-6] * 3) as sv:
        I, E = sv.solve_fp64_direct(s, INCS, (nt, np_), return_E=True)
    Iref, Eref = oracle.solve_direct(s.arrays(), [t.as_tuple() for t in INCS], nt, np_)
    for k in range(len(INCS)):
        assert rel(I[k], Iref[k]) < 1e-12
        assert rel(E[k].reshape(3, -1), Eref[k]) < 1e-12


def test_direct_fp64_device_inputs(solver):
    s = wf.fibonacci_sphere(1000, 1e-6)
    print("case: fp64 direct device tensors", flush=True)
    I = solver.solve_fp64_direct(_dev(s), INCS[:2], (12, 24))
    Iref, _ = oracle.solve_direct(s.arrays(), [t.as_tuple() for t in INCS[:2]], 12, 24)
    assert rel(I.cpu().numpy().reshape(2, -1), Iref) < 1e-12


@pytest.mark.parametrize("depth", [0, 1, 3, 4])
def test_octree_matches_reference_partition(solver, depth):
    s = _rand_surface(20000, 7)
    # points on split planes and outside the box exercise the float predicates
    s.x[:50] = 0.0
    s.y[50:100] = 0.0
    box = [-2e-6, -2e-6, -1.5e-6, 2e-6, 2e-6, 1.5e-6]
    print(f"case: octree depth={depth}", flush=True)
    with wf.POSolver(device=0, box=box) as sv:
        sv.set_geometry(s, depth)
        perm, nb, ne = [t.cpu().numpy() for t in sv.octree_nodes()]
    rperm, rb, re_, _ = oracle.octree([s.x, s.y, s.z], box, depth)
    np.testing.assert_array_equal(perm, rperm)
    np.testing.assert_array_equal(nb, rb)
    np.testing.assert_array_equal(ne, re_)


@pytest.fixture(params=[(0, 100000, 2), (0, 100000, 1), (0, 100000, 0), (0, 0, 2), (1, 288, 2)],
                ids=["agg_mfma_fused", "agg_mfma_fused1", "agg_mfma_2pass", "agg_fft2", "agg_fft"])
def agg_solver(request):
    """A solver with the level-transition kernels selected: dense split-BF16
    MFMA interpolation on every level (internal levels in the one-kernel fused
    form, staged (the default) or not, or as the phi and theta passes), the
    large-grid LDS FFT kernels on every level (5-smooth grids), or the
    original LDS FFT kernels."""
    import wavefiber3d_amd as wfm

    s = wfm.POSolver(device=0)
    s.tune = {"agg_kernel": request.param[0], "agg_fft_rows": request.param[1],
              "agg_fused": request.param[2]}
    yield s
    s.close()


@pytest.mark.parametrize("ntc,npc,ntp,npp,off", [
    (9, 16, 15, 28, None), (12, 20, 20, 36, None), (10, 18, 7, 12, "out"), (14, 26, 30, 60, "out"),
    (24, 48, 40, 80, None), (21, 35, 11, 22, "out"), (45, 80, 75, 144, None),
    (36, 60, 90, 180, "out"), (30, 50, 25, 40, None),
    # equal grids (the even length's Nyquist bin folds), a 240 -> 500 wave line,
    # theta lines above 1024 samples (the block-wide FFT kernels), and phi rows
    # of exactly twice the child's length
    (24, 48, 24, 48, None), (24, 48, 24, 48, "out"), (120, 240, 250, 500, None),
    (30, 60, 520, 1040, "out"), (20, 40, 36, 80, None), (60, 120, 100, 240, "out")])
def test_interp_matches_bruteforce(agg_solver, ntc, npc, ntp, npp, off):
    solver = agg_solver
    rng = np.random.default_rng(ntc * 100 + npp)
    # band-limited pattern: a few point sources radiating on the child grid
    off_c = -np.pi / 2 + np.pi / (2 * ntc)
    t = off_c + np.pi * np.arange(ntc) / ntc
    p = 2 * np.pi * np.arange(npc) / npc
    T, P = np.meshgrid(t, p, indexing="ij")
    sdir = np.stack([np.cos(T) * np.cos(P), np.cos(T) * np.sin(P), np.sin(T)], -1)
    pats = []
    for _ in range(3):
        src = rng.uniform(-1, 1, (4, 3)) * 1.2
        amp = rng.standard_normal(4) + 1j * rng.standard_normal(4)
        pats.append(np.einsum("s,ijs->ij", amp, np.exp(1j * np.einsum("ijc,sc->ijs", sdir, src))))
    child = np.stack(pats)
    off_p = -np.pi / 2 + 1e-10 if off == "out" else -np.pi / 2 + np.pi / (2 * ntp)
    print(f"case: interp {ntc}x{npc} -> {ntp}x{npp}", flush=True)
    out = solver.interp(torch.from_numpy(child).cuda(), ntp, npp, off_p).cpu().numpy()
    for k in range(child.shape[0]):
        ref = oracle.sphere_interp(child[k], ntp, npp, off_p)
        assert rel(out[k], ref) < 2e-5


@pytest.mark.parametrize("depth", [1, 2, 3])
def test_tree_sphere_vs_fp64(agg_solver, depth):
    solver = agg_solver
    s = wf.fibonacci_sphere(12000, 2e-6)
    nt, np_ = 90, 180
    print(f"case: tree sphere n=12000 depth={depth} agg {solver.tune['agg_kernel']}", flush=True)
    solver.set_geometry(s, depth)
    I = solver.solve(INCS, (nt, np_))
    Iref, _ = oracle.solve_direct(s.arrays(), [t.as_tuple() for t in INCS], nt, np_)
    for k in range(len(INCS)):
        assert rel(I[k], Iref[k]) < 1e-4, (k, rel(I[k], Iref[k]))


@pytest.mark.parametrize("depth", [2, 3])
def test_tree_root_reads_deeper_level(agg_solver, depth):
    """Output grid coarser than the level-1 grid: with the MFMA / LDS FFT
    kernels the root evaluates the first level whose grid fits the output grid
    directly (per-child translation table); the legacy kernels keep every
    level.  Both must match the FP64 direct sum."""
    solver = agg_solver
    s = wf.fibonacci_sphere(20000, 3e-6)
    nt, np_ = 20, 40
    print(f"case: root from a deeper level, depth={depth} agg {solver.tune['agg_kernel']}",
          flush=True)
    solver.set_geometry(s, depth)
    I = solver.solve(INCS, (nt, np_))
    lv = solver.stats()["levels"]
    assert lv[1][0] > nt or lv[1][1] > np_  # level 1 finer than the output grid
    Iref, _ = oracle.solve_direct(s.arrays(), [t.as_tuple() for t in INCS], nt, np_)
    for k in range(len(INCS)):
        assert rel(I[k], Iref[k]) < 1e-4, (k, rel(I[k], Iref[k]))


def test_tree_complex_E_vs_fp64(solver):
    s = wf.fibonacci_sphere(6000, 1.5e-6)
    nt, np_ = 36, 72
    print("case: tree complex E depth=2", flush=True)
    solver.set_geometry(s, 2)
    I, E = solver.solve(INCS[:2], (nt, np_), return_E=True)
    Iref, Eref = oracle.solve_direct(s.arrays(), [t.as_tuple() for t in INCS[:2]], nt, np_)
    for k in range(2):  # complex E carries the phase error as well: 2.4e-4 observed
        assert rel(E[k].reshape(3, -1), Eref[k]) < 5e-4


def test_tree_ellipsoid_fixture(agg_solver):
    solver = agg_solver
    import os

    g = os.path.join(os.path.dirname(__file__), "golden")
    s = wf.read_soa_mesh(os.path.join(g, "ellipsoid_3_3_2_float.binary"),
                         os.path.join(g, "ellipsoid_3_3_2_normal_float.binary"),
                         os.path.join(g, "ellipsoid_3_3_2_area_float.binary"))
    # compareBEM3d/ellipsoid.cu: lambda 500 nm, kappa 0.3, incidence +x, TM and TE
    incs = [wf.Incidence(0.0, 0.0, 500e-9, "M", 1.55 + 0.3j), wf.Incidence(0.0, 0.0, 500e-9, "E", 1.55 + 0.3j)]
    nt, np_ = 60, 120
    print("case: ellipsoid fixture depth=3", flush=True)
    solver.set_geometry(s, 3)
    I = solver.solve(incs, (nt, np_))
    Iref, _ = oracle.solve_direct(s.arrays(), [t.as_tuple() for t in incs], nt, np_)
    for k in range(2):
        assert rel(I[k], Iref[k]) < 1e-4, (k, rel(I[k], Iref[k]))


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "fp32_valu"])
def test_tree_precision_modes(precision):
    """Leaf GEMM on FP32 MFMA, split-BF16x3 MFMA and the vector ALU: each tree
    solve within 1e-4 of the FP64 oracle, and MFMA FP32 vs vector FP32 close."""
    s = wf.fibonacci_sphere(8000, 1.5e-6)
    # 7 identical solves share one wavelength group (F = 42 fields, 6 MFMA tiles)
    incs = INCS[:2] + [wf.Incidence(0.2, 0.4, 0.45e-6, "M", 1.55 + 0.05j)] * 7
    nt, np_ = 40, 80
    print(f"case: tree precision {precision} depth=3", flush=True)
    with wf.POSolver(device=0, precision=precision) as sv:
        sv.set_geometry(s, 3)
        I = sv.solve(incs, (nt, np_))
    Iref, _ = oracle.solve_direct(s.arrays(), [t.as_tuple() for t in incs[:3]], nt, np_)
    for k in range(3):
        assert rel(I[k], Iref[k]) < 1e-4, (precision, k, rel(I[k], Iref[k]))
    for k in range(3, len(incs)):
        assert rel(I[k], I[2]) < 1e-6  # identical incidences in other field slots


@pytest.mark.parametrize("chunk", [4, 8])
def test_quad_leaf_solves_per_pass(chunk):
    """Quad-symmetric leaf kernel with 4 or 8 solves per tree pass (8: 48 fields
    in 6 MFMA tiles; its epilogue projects two solves per lane group onto the
    transverse components): every solve within 1e-4 of the FP64 oracle."""
    s = wf.fibonacci_sphere(8000, 1.5e-6)
    incs = [wf.Incidence(0.3 * i - 1.0, 0.7 * i, 0.45e-6, "ME"[i % 2], 1.55 + 0.05j)
            for i in range(8)]
    nt, np_ = 40, 80
    print(f"case: quad leaf, {chunk} solves per pass, depth=3", flush=True)
    with wf.POSolver(device=0) as sv:
        sv.tune = {"leaf_quad": 1, "chunk_solves": chunk}
        sv.set_geometry(s, 3)
        I = sv.solve(incs, (nt, np_))
    Iref, _ = oracle.solve_direct(s.arrays(), [t.as_tuple() for t in incs], nt, np_)
    for k in range(len(incs)):
        assert rel(I[k], Iref[k]) < 1e-4, (chunk, k, rel(I[k], Iref[k]))


@pytest.mark.parametrize("chunk", [2, 4, 8])
def test_quad_leaf_groups_per_block(chunk):
    """Quad leaf kernel with all of a leaf's groups of 64 direction quads in one
    block (leaf_quad 3; the 12 x 24 leaf grid has two groups) against one group
    per block (leaf_quad 2): bit-identical (each group runs the same operations
    on the same staged values), and within 1e-4 of the FP64 oracle.  Leaves of
    up to one staging chunk of points are staged once for both groups, larger
    ones per group; 2 solves per pass (2 tiles) take the epilogue through the
    LDS and restage per group, 4 and 8 (3 and 6 tiles) project in registers."""
    s = wf.fibonacci_sphere(8000, 1.5e-6)
    incs = [wf.Incidence(0.3 * i - 1.0, 0.7 * i, 0.45e-6, "ME"[i % 2], 1.55 + 0.05j)
            for i in range(8)]
    nt, np_ = 40, 80
    out = {}
    for mode in (2, 3):
        print(f"case: quad leaf mode {mode}, {chunk} solves per pass, depth=3", flush=True)
        with wf.POSolver(device=0) as sv:
            sv.tune = {"leaf_quad": mode, "chunk_solves": chunk}
            sv.set_geometry(s, 3)
            out[mode] = sv.solve(incs, (nt, np_))
    assert np.array_equal(out[3], out[2]), rel(out[3], out[2])
    Iref, _ = oracle.solve_direct(s.arrays(), [t.as_tuple() for t in incs], nt, np_)
    for k in range(len(incs)):
        assert rel(out[3][k], Iref[k]) < 1e-4, (chunk, k, rel(out[3][k], Iref[k]))


@pytest.mark.parametrize("version", [1, 2])
@pytest.mark.parametrize("depth", [3, 4])
def test_fused_transition_matches_two_pass(depth, version):
    """The one-kernel internal level transitions (phi output in LDS; version
    2 also stages the phi matrix and the child rows in LDS) and the two-pass
    MFMA transition use the same split-BF16 matrices: their trees agree to
    FP32 rounding, both meet the oracle, and the fused form runs no separate
    phi pass below the root."""
    s = wf.fibonacci_sphere(60000, 4e-6)
    nt, np_ = 60, 120
    incs = INCS + [wf.Incidence(0.1 * k, 0.3 * k, 0.5e-6, "ME"[k % 2], 1.5 + 0.05j) for k in range(4)]
    print(f"case: fused v{version} vs two-pass transitions, sphere n=60000 depth={depth}",
          flush=True)
    out = {}
    for fused in (version, 0):
        with wf.POSolver(device=0, solve_streams=1) as sv:
            sv.tune = {"agg_fused": fused, "agg_fft_rows": 100000}
            sv.set_geometry(s, depth)
            out[fused] = (sv.solve(incs, (nt, np_)), sv.stats())
    Iref, _ = oracle.solve_direct(s.arrays(), [t.as_tuple() for t in incs], nt, np_)
    for k in range(len(incs)):
        e = rel(out[version][0][k], out[0][0][k])
        assert e < 2e-6, (k, e)
        assert rel(out[version][0][k], Iref[k]) < 1e-4
    phi_fused = out[version][1]["kernels"].get("phi_pass", {"launches": 0})["launches"]
    phi_2pass = out[0][1]["kernels"]["phi_pass"]["launches"]
    assert phi_fused < phi_2pass
