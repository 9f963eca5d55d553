"""Automatic depth, autotuning, hipGraph replay and batched multi-segment
solves on the GPU.

* A WF_DEPTH_AUTO geometry solved at depth d (chosen by the cost model, or
  forced through autotuning) agrees with the depth-d build of the same
  surface: the truncated tree has the same level lists, only the order of
  points inside a leaf differs (1e-5 relative L2), and with the FP64 direct
  sum to the tree gate (1e-3).
* Graph replays: every replay equals the eager solve bit for bit, with other
  eager work between the replays.
* wf_batch_*: every segment equals the FP64 direct sum of that segment alone
  (1e-3) and fixed-depth segments equal a single-handle solve (1e-5); a
  second batch solve (graph replays) repeats the first bit for bit; device
  and host outputs agree.
"""
import numpy as np
import pytest

import wavefiber3d_amd as wf

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

GRID = (32, 64)
INCS = [
    wf.Incidence(np.pi / 2 - 0.3, 0.4, 0.5e-6, "M", 1.55 + 0.02j),
    wf.Incidence(0.2, 1.1, 0.5e-6, "E", 1.55 + 0.02j),
    wf.Incidence(-0.5, 2.0, 0.62e-6, "M", 1.55 + 0.1j),
]


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _fp64(surface, incs=INCS, grid=GRID):
    with wf.POSolver(device=0) as sv:
        return sv.solve_fp64_direct(surface, incs, grid)


def _ellipsoid(n, a, b, c, seed):
    s = wf.fibonacci_sphere(n, 1.0)
    x, y, z, nx, ny, nz, ar = s.arrays()
    X, Y, Z = a * x, b * y, c * z
    N = np.stack([nx / a, ny / b, nz / c])
    N /= np.linalg.norm(N, axis=0)
    rng = np.random.default_rng(seed)
    area = ar * (a * b * c) ** (2 / 3) * rng.uniform(0.9, 1.1, n)
    return wf.SurfaceSamples(*[v.astype(np.float32) for v in (X, Y, Z, N[0], N[1], N[2], area)])


def test_auto_depth_matches_fixed_build_and_fp64():
    s = wf.fibonacci_sphere(20000, 2.5e-6)
    print("case: auto depth sphere n=20000", flush=True)
    with wf.POSolver(device=0, precision="fp32") as sv:
        sv.set_geometry(s, "auto")
        assert sv.depth == 5  # ceil(log8 20000)
        I_auto = sv.solve(INCS, GRID)
        d = sv.stats()["depth"]
        assert 1 <= d <= 5
        sv.set_geometry(s, d)
        I_fix = sv.solve(INCS, GRID)
    r_fix = rel(I_auto, I_fix)
    r_64 = rel(I_auto, _fp64(s))
    print(f"  chosen depth {d}: vs fixed build {r_fix:.2e}, vs fp64 {r_64:.2e}", flush=True)
    assert r_fix < 1e-5
    assert r_64 < 1e-3


def test_autotune_calibrates_and_keeps_accuracy():
    s = _ellipsoid(30000, 3e-6, 2e-6, 1.5e-6, 1)
    print("case: autotune ellipsoid n=30000", flush=True)
    with wf.POSolver(device=0) as sv:
        sv.set_geometry(s, "auto")
        best = sv.autotune(INCS, GRID)
        t = sv.tune
        print(f"  best depth {best}, tune {t}", flush=True)
        assert 1 <= best <= sv.depth
        assert t["calibrated"] == 1
        assert t["leaf_s_per_pair"] > 0 and t["agg_s_per_byte"] > 0 and t["level_s"] >= 0
        assert t["chunk_solves"] in (4, 8)
        assert (t["agg_fused"], t["theta_tile"]) in ((2, 0), (0, 0), (2, 2))
        I = sv.solve(INCS, GRID)
        # a calibration moves between handles
        with wf.POSolver(device=0) as sv2:
            sv2.tune = t
            assert sv2.tune == t
    assert rel(I, _fp64(s)) < 1e-3


def test_graph_replays_equal_eager():
    s = wf.fibonacci_sphere(15000, 2e-6)
    dev = [torch.from_numpy(a).cuda() for a in s.arrays()]
    print("case: graph replay sphere n=15000 depth 3", flush=True)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        with wf.POSolver(device=0, use_graph=True) as sv:
            sv.set_geometry(dev, 3)
            I0 = sv.solve(INCS, GRID, out="torch").clone()  # eager + capture
            for rep in range(3):
                x = torch.randn(1 << 20, device="cuda")  # eager work between replays
                y = (x * 2).sum()
                I = sv.solve(INCS, GRID, out="torch")
                stream.synchronize()
                assert torch.equal(I, I0), f"replay {rep}"
                del x, y
            # new incidences with the same signature shape -> re-captured, still right
            incs2 = [wf.Incidence(0.3, 0.2, 0.5e-6, "M", 1.55 + 0.02j)] + INCS[1:]
            I2 = sv.solve(incs2, GRID, out="torch")
            stream.synchronize()
    with wf.POSolver(device=0) as ref:
        ref.set_geometry(s, 3)
        J0 = ref.solve(INCS, GRID)
        J2 = ref.solve(incs2, GRID)
    assert np.array_equal(I0.cpu().numpy(), J0)
    assert np.array_equal(I2.cpu().numpy(), J2)


def _segments():
    segs = []
    for i, (n, a, b, c) in enumerate([(24000, 3e-6, 2.5e-6, 2e-6), (6000, 1.5e-6, 1.5e-6, 1e-6),
                                      (15000, 2e-6, 3e-6, 1.5e-6), (3000, 1e-6, 1e-6, 1e-6),
                                      (9000, 2.5e-6, 1.5e-6, 1.5e-6)]):
        segs.append(_ellipsoid(n, a, b, c, 10 + i))
    fib = wf.fiber_segment(r1=3e-6, r2=2e-6, length=1e-6, spacing=0.05e-6, seed=3)
    segs.append(fib)
    return segs


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_single_solves(use_graph):
    segs = _segments()
    depths = ["auto", 2, "auto", 1, 3, "auto"]
    print(f"case: batch of {len(segs)} segments use_graph={use_graph}", flush=True)
    # beta = 2 is the autotuner's floor, so the batch's calibration (on its
    # largest auto-depth segment) keeps the sampling of the single-handle
    # solves below; its kernel choices change results by FP32 rounding only
    with wf.Batch(n_streams=3, device=0, use_graph=use_graph, beta=2.0) as b:
        for s, d in zip(segs, depths):
            b.add(s, d)
        I = b.solve(INCS, GRID)
        info = b.info()
        I2 = b.solve(INCS, GRID)
        Id = b.solve(INCS, GRID, out="torch").cpu().numpy()
    print(f"  depths {info['depth'].tolist()} ms {np.round(info['ms'], 2).tolist()}", flush=True)
    assert I.shape == (len(segs), len(INCS)) + GRID
    assert np.array_equal(I, I2)
    assert np.array_equal(I, Id)
    assert (info["ms"] > 0).all()
    for k, (s, d) in enumerate(zip(segs, depths)):
        r = rel(I[k], _fp64(s))
        print(f"  segment {k} n={s.n} depth {d}->{info['depth'][k]}: vs fp64 {r:.2e}", flush=True)
        assert r < 1e-3
        if d != "auto":
            assert info["depth"][k] == d
            with wf.POSolver(device=0, beta=2.0) as sv:
                sv.tune = {"chunk_solves": 4, "leaf_quad": 1}
                sv.set_geometry(s, d)
                J = sv.solve(INCS, GRID)
            assert rel(I[k], J) < 1e-5


def test_batch_rejects_bad_grid_and_empty():
    with wf.Batch(n_streams=1, device=0) as b:
        with pytest.raises(wf.WfError):
            b.solve(INCS, GRID)  # no segments
        b.add(wf.fibonacci_sphere(2000, 1e-6), 2)
        with pytest.raises(wf.WfError):
            b.solve(INCS, (32, 63))  # odd n_phi


@pytest.mark.parametrize("depth", [0, 3])
@pytest.mark.parametrize("field_chunk", [0, 12])
def test_solve_streams_identical_to_sequential(field_chunk, depth):
    """wf_solve's concurrent chunks (opts.solve_streams) give bit-identical
    results to the sequential order, on device and host outputs, for tree and
    depth-0 (FP64 direct) geometries, and the same launch counts in the merged
    statistics."""
    s = wf.fibonacci_sphere(20000, 2e-6)
    dev = [torch.from_numpy(a).cuda() for a in s.arrays()]
    incs = [wf.Incidence(0.3 * k, 0.5, 0.45e-6 + 0.02e-6 * (k // 2), "ME"[k % 2], 1.5 + 0.01j)
            for k in range(12)]
    grid = (36, 72)
    out = {}
    for K in (1, 3, 4):
        print(f"case: solve_streams={K} field_chunk={field_chunk} depth={depth}", flush=True)
        with wf.POSolver(device=0, solve_streams=K, field_chunk=field_chunk) as sv:
            sv.set_geometry(dev, depth)
            I = sv.solve(incs, grid, out="torch")
            torch.cuda.synchronize()
            st = sv.stats()
            Ih = sv.solve(incs, grid)
            out[K] = (I.cpu().numpy(), Ih, {k: v["launches"] for k, v in st["kernels"].items()})
    for K in (3, 4):
        assert np.array_equal(out[K][0], out[1][0])
        assert np.array_equal(out[K][1], out[1][1])
        assert out[K][2] == out[1][2]


def test_solve_streams_cancel_then_resume():
    """A cancellation flag set by the host stops a multi-stream solve with
    WF_ERR_CANCELLED (every worker stops at its next chunk, the caller's stream
    still joins them); cleared, the next solve on the same handle matches a
    sequential handle bit for bit, complex E included."""
    import ctypes as ct

    s = wf.fibonacci_sphere(12000, 1.5e-6)
    dev = [torch.from_numpy(a).cuda() for a in s.arrays()]
    incs = [wf.Incidence(0.2 * k, 0.1, 0.5e-6 + 0.03e-6 * (k // 2), "ME"[k % 2], 1.5 + 0.0j)
            for k in range(8)]
    grid = (30, 60)
    flag = ct.c_int32(1)
    print("case: cancelled multi-stream solve", flush=True)
    with wf.POSolver(device=0, solve_streams=4, cancel_flag=flag) as sv:
        sv.set_geometry(dev, 2)
        with pytest.raises(wf.WfError, match="CANCELLED"):
            sv.solve(incs, grid, out="torch")
        torch.cuda.synchronize()
        flag.value = 0
        I, E = sv.solve(incs, grid, return_E=True, out="torch")
        I, E = I.cpu().numpy(), E.cpu().numpy()
    with wf.POSolver(device=0, solve_streams=1) as sq:
        sq.set_geometry(dev, 2)
        I1, E1 = sq.solve(incs, grid, return_E=True, out="torch")
        I1, E1 = I1.cpu().numpy(), E1.cpu().numpy()
    assert np.array_equal(I, I1)
    assert np.array_equal(E, E1)


def test_batch_cancel_then_resume():
    """A cancelled batch solve returns WF_ERR_CANCELLED after joining every
    worker stream back into the caller's stream (device output); cleared, the
    next batch solve on the same batch matches a fresh batch bit for bit."""
    import ctypes as ct

    segs = [wf.fibonacci_sphere(n, r) for n, r in ((9000, 1.5e-6), (5000, 1.2e-6), (3000, 1e-6))]
    flag = ct.c_int32(1)
    print("case: cancelled batch solve, 3 segments, 3 streams", flush=True)
    with wf.Batch(n_streams=3, device=0, cancel_flag=flag) as b:
        for s in segs:
            b.add(s, 2)
        with pytest.raises(wf.WfError, match="CANCELLED"):
            b.solve(INCS, GRID, out="torch")
        torch.cuda.synchronize()
        flag.value = 0
        I = b.solve(INCS, GRID, out="torch").cpu().numpy()
    with wf.Batch(n_streams=3, device=0) as b1:
        for s in segs:
            b1.add(s, 2)
        I1 = b1.solve(INCS, GRID)
    assert np.array_equal(I, I1)
