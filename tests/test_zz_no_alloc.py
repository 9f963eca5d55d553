"""The allocation-free 2D cross-section entry point (wf_rough2d_solve_noalloc)
against wf_rough2d_solve through the C-ABI: two calls in a row, the second
with a smaller input (fewer samples, incidences and azimuths), each equal to
the allocating entry point bit for bit (same kernels, same split of the
sample range), and the size limit reported as an error."""
import numpy as np
import pytest

import wavefiber3d_amd as wf  # noqa: F401  (loads the library)
from wavefiber3d_amd import rough2d as r2
from wavefiber3d_amd._lib import WfError

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _ellipse(n, a, b):
    t = 2 * np.pi * np.arange(n) / n
    return a * np.cos(t), b * np.sin(t)


def test_noalloc_matches_allocating_entry_point_twice():
    cases = [(2000, 25, 3600, "M"), (700, 3, 900, "E")]
    for n, n_inc, n_phi, mode in cases:
        print(f"case: rough2d noalloc n={n} n_inc={n_inc} n_phi={n_phi} {mode}", flush=True)
        x, y = _ellipse(n, 12e-6, 8e-6)
        incs = [r2.Rough2dIncidence(r2.wavelength(l), mode, theta=0.05 * (l % 3))
                for l in range(n_inc)]
        xd = torch.from_numpy(x).cuda()
        yd = torch.from_numpy(y).cuda()
        ref = r2.rough2d_solve(xd, yd, incs, n_phi=n_phi)
        got = r2.rough2d_solve_noalloc(xd, yd, incs, n_phi=n_phi)
        torch.cuda.synchronize()
        assert torch.equal(got, ref)


def test_noalloc_rejects_sizes_beyond_the_scratch():
    x, y = _ellipse(20000, 12e-6, 8e-6)
    incs = [r2.Rough2dIncidence(r2.wavelength(l), "M") for l in range(25)]
    xd = torch.from_numpy(x).cuda()
    yd = torch.from_numpy(y).cuda()
    print("case: rough2d noalloc beyond the scratch size", flush=True)
    with pytest.raises(WfError, match="scratch"):
        r2.rough2d_solve_noalloc(xd, yd, incs, n_phi=3600)
    with pytest.raises(WfError, match="n_inc"):
        r2.rough2d_solve_noalloc(xd[:100], yd[:100],
                                 [r2.Rough2dIncidence(0.5e-6, "M")] * 1025, n_phi=8)
