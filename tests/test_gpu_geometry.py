"""Geometry front-end on the GPU (wf_heightfield_points): bit-identical to the
float32 oracle (util/geometry.h:78-178) for host and device buffers, arc
windows including the wrap at the last arc sample, the last ring, tiny
grids; a zero vertex is reported at the first sample of the (i, j) loop."""
import numpy as np
import pytest

import wavefiber3d_amd as wf
from oracle import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _fiber(r1, r2, length, spacing, seed):
    hx, hy, nv = wf.geometry.fiber_heightfield(r1, r2, length, spacing, seed=seed)
    h = np.concatenate([hx.astype(np.float32).ravel(), hy.astype(np.float32).ravel()])
    return h, hx.shape[0], nv


CASES = [
    ("fiber 10x8um", (10e-6, 8e-6, 3e-6, 0.1e-6, 1), [(0, None), (5, 100), (-40, None)]),
    ("fiber 40x30um, 663k", (40e-6, 30e-6, 30e-6, 0.1e-6, 0), [(0, None), (552, 1657)]),
    ("tiny", (1e-6, 1e-6, 1e-6, 0.5e-6, 2), [(0, None), (1, 1)]),
]


@pytest.mark.parametrize("name, geo, windows", CASES)
def test_heightfield_bit_identical_to_oracle(name, geo, windows):
    h, numel, nv = _fiber(*geo)
    period = geo[2]
    unitz = np.float32(period / nv)
    hd = torch.from_numpy(h).cuda()
    for s, e in windows:
        s = s % numel
        e_ = numel - 1 if e is None else e
        print(f"case: heightfield {name} numel={numel} numelvert={nv} window=[{s}, {e_}]",
              flush=True)
        ref, bad = oracle.heightfield_points(h, numel, nv, period, unitz, s, e_)
        assert bad == -1
        host = wf.heightfield_samples(h, numel, nv, period, float(unitz), s, e_).arrays()
        dev = [t.cpu().numpy() for t in
               wf.heightfield_samples(hd, numel, nv, period, float(unitz), s, e_)]
        torch.cuda.synchronize()
        for c in range(7):
            assert np.array_equal(host[c], ref[c]), (name, c)
            assert np.array_equal(dev[c], ref[c]), (name, c)


def test_heightfield_zero_vertex_error():
    numel, nv = 16, 5
    phi = 2 * np.pi * np.arange(numel) / numel
    hx = np.repeat((4e-6 * np.cos(phi))[:, None], nv, 1).astype(np.float32)
    hy = np.repeat((4e-6 * np.sin(phi))[:, None], nv, 1).astype(np.float32)
    hx[3, 2] = hy[3, 2] = 0.0
    h = np.concatenate([hx.ravel(), hy.ravel()])
    print("case: heightfield zero vertex", flush=True)
    _, bad = oracle.heightfield_points(h, numel, nv, 1e-6, np.float32(0.2e-6))
    assert bad == 1 * numel + 2
    with pytest.raises(wf.WfError, match="j = 2, i = 1"):
        wf.heightfield_samples(h, numel, nv, 1e-6, 0.2e-6)
    with pytest.raises(wf.WfError, match="j = 2, i = 1"):
        wf.heightfield_samples(torch.from_numpy(h).cuda(), numel, nv, 1e-6, 0.2e-6)
