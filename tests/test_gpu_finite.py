"""Finite-distance FP64 validator on the GPU (wf_po_solve_fp64_finite;
compareBEM3d/ellipsoid_simple.cu:229-261, :384-589).

* The reference's validator configuration -- the 30 322-sample ellipsoid mesh,
  500 nm, eta 1.55 + 0.3i, incidence along +x, TM, beam sigma 1 cm, the
  theta = 0 ring at D = 1 m -- against the oracle's finite sum (intensity and
  projected field), device and host inputs, several incidences at once.
* Far away the finite solve times D^2 equals the far-field FP64 solve.
* A distance that is not positive and finite is rejected.
"""
import os

import numpy as np
import pytest

import wavefiber3d_amd as wf
from oracle import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _mesh():
    g = os.path.join(os.path.dirname(__file__), "golden")
    return wf.read_soa_mesh(os.path.join(g, "ellipsoid_3_3_2_float.binary"),
                            os.path.join(g, "ellipsoid_3_3_2_normal_float.binary"),
                            os.path.join(g, "ellipsoid_3_3_2_area_float.binary"))


def rel(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# ellipsoid_simple.cu:398-399, 470-482, 492: sigma 1 cm, 500 nm, kappa 0.3, TM
INC = wf.Incidence(0.0, 0.0, 500e-9, "M", 1.55 + 0.3j, 1e-2)


@pytest.mark.parametrize("device_input", [True, False])
def test_finite_ring_matches_oracle(solver, device_input):
    s = _mesh()
    # theta index 1 of 2 is the ring theta = 1e-10 (the reference uses 0)
    nt, np_ = 2, 720
    print(f"case: finite ellipsoid D=1 grid={nt}x{np_} device_input={device_input}", flush=True)
    surf = [torch.from_numpy(a).cuda() for a in s.arrays()] if device_input else s
    I, E = solver.solve_fp64_finite(surf, [INC], (nt, np_), 1.0, return_E=True)
    if device_input:
        I, E = I.cpu().numpy(), E.cpu().numpy()
    Iref, Eref = oracle.solve_finite(s.arrays(), [INC.as_tuple()], nt, np_, 1.0)
    assert rel(I[0], Iref[0]) < 1e-9
    assert rel(E[0].reshape(3, -1), Eref[0]) < 1e-9


def test_finite_several_incidences(solver):
    s = _mesh()
    incs = [wf.Incidence(0.0, 0.0, 500e-9, "M", 1.55 + 0.3j, 1e-2),
            wf.Incidence(0.4, 1.0, 500e-9, "E", 1.55 + 0.3j, 0.0),
            wf.Incidence(-0.3, 2.0, 620e-9, "M", 1.4 + 0.0j, 5e-6)]
    nt, np_ = 12, 40
    print(f"case: finite ellipsoid D=0.25 grid={nt}x{np_} 3 incidences", flush=True)
    I = solver.solve_fp64_finite(s, incs, (nt, np_), 0.25)
    Iref, _ = oracle.solve_finite(s.arrays(), [t.as_tuple() for t in incs], nt, np_, 0.25)
    for k in range(3):
        assert rel(I[k], Iref[k]) < 1e-9, k


def test_finite_far_limit_equals_far_field(solver):
    s = _mesh()
    incs = [INC, wf.Incidence(0.2, 0.5, 500e-9, "E", 1.55 + 0.3j, 1e-2)]
    nt, np_ = 30, 60
    D = 1e3
    print(f"case: finite ellipsoid D={D:g} vs far field grid={nt}x{np_}", flush=True)
    If = solver.solve_fp64_finite(s, incs, (nt, np_), D)
    Ifar = solver.solve_fp64_direct(s, incs, (nt, np_))
    # the plane-wave phase error k |x|^2 / (2D) is below 1e-7 rad here
    for k in range(2):
        assert rel(If[k] * D * D, Ifar[k]) < 1e-6, k
    # at 1 m the curvature of the wavefront still shows (k |x|^2 / 2 ~ 6e-5 rad)
    I1 = solver.solve_fp64_finite(s, incs[:1], (nt, np_), 1.0)
    assert 1e-7 < rel(I1[0], Ifar[0]) < 1e-2


@pytest.mark.parametrize("bad", [0.0, -1.0, float("inf"), float("nan")])
def test_finite_rejects_bad_distance(solver, bad):
    s = _mesh()
    with pytest.raises(wf.WfError, match="distance"):
        solver.solve_fp64_finite(s, [INC], (2, 8), bad)
