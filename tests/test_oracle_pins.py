"""Pins of the FP64 oracle (CPU, no GPU needed).

The reference ships no golden vectors or known-answer tests for the PO path,
so each oracle function is pinned by something the mathematics fixes:
textbook Fresnel forms, the flat-plate specular closed form (Fresnel
reflection of a plane wave, which ties currents + radiation factors + direct
sum + intensity + polarization mapping together), Dirichlet-kernel lattice
sums, dipole patterns, band-limited interpolation exactness, and an
independent sort-based restatement of the octree partition.
"""
import numpy as np
import pytest

import oracle as O

LAM = 0.5e-6
K = 2 * np.pi / LAM


# ------------------------------------------------------------- Fresnel -----
@pytest.mark.parametrize("eta", [1.55, 1.55 + 0.1j, 1.3 + 0.02j, 2.4 + 1.0j])
def test_fresnel_textbook(eta):
    for th in np.linspace(0, np.pi / 2 - 1e-6, 23):
        rs, rp = O.fresnel(th, eta)
        c = np.cos(th)
        st = np.sqrt(eta * eta - np.sin(th) ** 2 + 0j)  # = eta cos(theta_t)
        rs_t = (c - st) / (c + st)
        rp_t = (eta * eta * c - st) / (eta * eta * c + st)
        # the reference form sqrt(1 - sin^2) cancels near grazing: ~1e-16/cos error
        tol = 1e-14 + 1e-15 / max(c, 1e-16)
        assert abs(rs - rs_t) < tol
        assert abs(rp - rp_t) < tol
        assert abs(rs) <= 1 + 1e-12 and abs(rp) <= 1 + 1e-12


def test_fresnel_special_angles():
    eta = 1.55
    rs, rp = O.fresnel(0.0, eta)
    assert abs(rs - (1 - eta) / (1 + eta)) < 1e-15
    assert abs(rp + rs) < 1e-15
    _, rpb = O.fresnel(np.arctan(eta), eta)  # Brewster
    assert abs(rpb) < 1e-14
    rs, rp = O.fresnel(np.pi / 2, eta)  # grazing
    assert abs(rs + 1) < 1e-12 and abs(rp + 1) < 1e-12


# ---------------------------------------------- incidence and currents -----
def test_incidence_basis_orthonormal():
    for th, ph in [(0.3, 1.0), (-1.2, 4.0), (np.pi / 2, 0.0)]:
        for pol in "ME":
            w, E, H = O.incidence_basis(th, ph, pol)
            assert abs(np.linalg.norm(w) - 1) < 1e-15
            assert abs(w @ E) < 1e-15 and abs(w @ H) < 1e-15 and abs(E @ H) < 1e-15
            # propagation along E x H (TM) or along H x E... right-handed set up to sign
            assert abs(abs(np.cross(E, H) @ w) - 1) < 1e-14


def test_radiation_factors():
    ef, hf = O.radiation_factors(LAM)
    mu0, eps0, c0 = 4e-7 * np.pi, 8.8541878e-12, 299792458.0
    w = K * c0
    g = np.exp(-1j * K)
    assert abs(ef - (-1j * w * mu0 / (4 * np.pi) * g)) < 1e-12 * abs(ef)
    assert abs(hf - (-1j * w * eps0 / (4 * np.pi) * g)) < 1e-12 * abs(hf)


def _plate(m=48, h=LAM / 8):
    g = (np.arange(m) - (m - 1) / 2) * h
    X, Y = np.meshgrid(g, g, indexing="ij")
    X, Y = X.ravel(), Y.ravel()
    n = X.size
    return [X, Y, np.zeros(n), np.zeros(n), np.zeros(n), np.ones(n), np.full(n, h * h)], (m * h) ** 2


@pytest.mark.parametrize("ti_deg", [-20.0, -45.0, -70.0])
@pytest.mark.parametrize("pol", ["M", "E"])
@pytest.mark.parametrize("eta", [1.55, 1.55 + 0.3j])
def test_flat_plate_specular_closed_form(ti_deg, pol, eta):
    """A plate reflects a plane wave into the specular direction with amplitude
    k A cos(theta_i) |r| / (2 pi D) (Fresnel r for the s / p part)."""
    surf, A = _plate()
    ti, phi = np.radians(ti_deg), 0.7
    w, E, _ = O.incidence_basis(ti, phi, pol)
    s = w - 2 * w[2] * np.array([0, 0, 1.0])
    I, _ = O.solve_direct(surf, [(ti, phi, LAM, pol, eta, 0.0)], 0, 0, dirs=s[None])
    rs, rp = O.fresnel(np.arccos(abs(w[2])), eta)
    pn = np.cross(w, [0, 0, 1.0])
    pn /= np.linalg.norm(pn)
    fs = (E @ pn) ** 2
    amp2 = (K * A * abs(w[2]) / (2 * np.pi)) ** 2
    pred = amp2 * (fs * abs(rs) ** 2 + (1 - fs) * abs(rp) ** 2)
    # the reference's c0, mu0, eps0 are mutually consistent only to ~1e-8
    assert abs(I[0, 0] / pred - 1) < 1e-7


def test_flat_plate_brewster_tm_zero():
    surf, A = _plate()
    eta = 1.55
    ti = -(np.pi / 2 - np.arctan(eta))  # local angle = Brewster
    w, _, _ = O.incidence_basis(ti, 0.0, "M")
    s = w - 2 * w[2] * np.array([0, 0, 1.0])
    I, _ = O.solve_direct(surf, [(ti, 0.0, LAM, "M", eta, 0.0)], 0, 0, dirs=s[None])
    Ie, _ = O.solve_direct(surf, [(ti, 0.0, LAM, "E", eta, 0.0)], 0, 0, dirs=s[None])
    assert I[0, 0] < 1e-20 * Ie[0, 0]


def _rand_surf(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.uniform(-2e-6, 2e-6, (3, n))
    q = rng.standard_normal((3, n))
    q /= np.linalg.norm(q, axis=0)
    return [p[0], p[1], p[2], q[0], q[1], q[2], rng.uniform(1e-15, 2e-15, n)]


def test_currents_tangential_and_shadow():
    surf = _rand_surf(500, 3)
    nrm = np.stack(surf[3:6], 1).astype(np.float32).astype(np.float64)
    for inc in [(0.3, 1.0, LAM, "M", 1.55 + 0.1j, 0.0), (-0.9, 2.0, LAM, "E", 1.3, 0.0)]:
        C = O.currents(surf, inc)
        w, _, _ = O.incidence_basis(inc[0], inc[1], inc[3])
        J, M = C[:, :3], C[:, 3:]
        assert np.abs(np.einsum("nc,nc->n", J, nrm)).max() < 1e-12 * np.abs(J).max()
        assert np.abs(np.einsum("nc,nc->n", M, nrm)).max() < 1e-12 * np.abs(M).max()
        shadow = nrm @ w > 0
        assert shadow.any() and (~shadow).any()
        assert np.all(C[shadow] == 0)


@pytest.mark.parametrize("pol", ["M", "E"])
def test_normal_incidence_branch(pol):
    """tangentplane.h:40-49: at w.n = -1 the reference scales both E and H by
    R = (1-eta)/(1+eta) (TM) or (eta-1)/(1+eta) (TE) -- quirk Q3 kept as is."""
    eta = 1.55 + 0.1j
    surf = [np.array([1e-7]), np.array([-2e-7]), np.array([3e-7]), np.zeros(1), np.zeros(1),
            -np.ones(1), np.array([1e-15])]
    ti = np.pi / 2  # w = (~0, 0, 1)
    C = O.currents(surf, (ti, 0.0, LAM, pol, eta, 0.0))[0]
    w, E, H = O.incidence_basis(ti, 0.0, pol)
    z0 = O.z0()
    x = np.array([1e-7, -2e-7, 3e-7], dtype=np.float32).astype(np.float64)
    e0 = np.exp(-1j * K * (w @ x))
    R = (1 - eta) / (1 + eta) if pol == "M" else (eta - 1) / (1 + eta)
    n = np.array([0, 0, -1.0])
    Et = E * e0 * (1 + R)
    Ht = H * e0 / z0 * (1 + R)
    ef, hf = O.radiation_factors(LAM)
    a = float(np.float32(1e-15))  # areas are float32 at the boundary
    J = np.cross(n, Ht) * a * ef
    M = np.cross(Et, n) * a * hf
    assert np.abs(C[:3] - J).max() < 1e-13 * np.abs(J).max()
    assert np.abs(C[3:] - M).max() < 1e-13 * np.abs(M).max()


def test_beam_taper():
    surf = _rand_surf(200, 4)
    inc0 = (0.2, 0.5, LAM, "M", 1.55, 0.0)
    sig = 1.5e-6
    C0 = O.currents(surf, inc0)
    C1 = O.currents(surf, inc0[:5] + (sig,))
    z = np.asarray(surf[2], np.float32).astype(np.float64)
    g = np.exp(-0.5 * z * z / sig ** 2)
    assert np.abs(C1 - C0 * g[:, None]).max() < 1e-13 * np.abs(C0).max()


# ---------------------------------------------------------- direct sum -----
def test_direct_sum_dirichlet_lattice():
    """m x m lattice, constant current: F = C D_m(k s_x h) D_m(k s_y h)."""
    m, h = 11, 0.37e-6
    g = (np.arange(m) - (m - 1) / 2) * h
    X, Y = np.meshgrid(g, g, indexing="ij")
    xyz = [X.ravel(), Y.ravel(), np.zeros(m * m)]
    c = np.array([1 + 2j, -0.5j, 0.3, 2.0, 1j, -1.0])
    C = np.tile(c, (m * m, 1))[None]
    dirs = O.output_dirs(13, 26)
    F = O.direct_sum(xyz, C, K, dirs)[0]
    hh = np.float32(h) * 0 + np.asarray(g, np.float32).astype(np.float64)  # lattice as stored
    ux, uy = K * dirs[:, 0], K * dirs[:, 1]
    Dx = np.exp(1j * np.outer(ux, hh)).sum(1)
    Dy = np.exp(1j * np.outer(uy, hh)).sum(1)
    # closed form of the centred lattice sum (exact float coordinates are symmetric)
    def dirich(u):
        u = np.where(np.abs(u) < 1e-12, 1e-12, u)
        return np.sin(m * u * h / 2) / np.sin(u * h / 2)
    assert np.abs(Dx - dirich(ux)).max() < 1e-6 * m
    F_ref = (Dx * Dy)[:, None] * c[None]
    assert np.abs(F - F_ref).max() < 1e-10 * np.abs(F_ref).max()


def test_direct_sum_translation_identity():
    surf = _rand_surf(300, 5)
    C = O.currents(surf, (0.1, 0.2, LAM, "E", 1.55, 0.0))[None]
    dirs = O.output_dirs(7, 14)
    c1 = np.array([1e-7, -3e-7, 2e-7])
    F0 = O.direct_sum(surf[:3], C, K, dirs, centre=(0, 0, 0))[0]
    F1 = O.direct_sum(surf[:3], C, K, dirs, centre=c1)[0]
    ph = np.exp(1j * K * dirs @ (0 - c1))
    assert np.abs(F1 - F0 * ph[:, None]).max() < 1e-12 * np.abs(F0).max()


# ----------------------------------------------------------- intensity -----
def test_intensity_dipoles():
    dirs = O.output_dirs(9, 18)
    z0 = O.z0()
    Fj = np.zeros((dirs.shape[0], 6), complex)
    Fj[:, 2] = 2.0 - 1j  # electric dipole along z
    I, E = O.intensity(dirs, Fj)
    sin2 = 1 - dirs[:, 2] ** 2
    assert np.allclose(I, abs(2 - 1j) ** 2 * sin2, rtol=1e-13, atol=1e-15)
    assert np.abs(np.einsum("cd,dc->d", E, dirs)).max() < 1e-14
    Fm = np.zeros_like(Fj)
    Fm[:, 3] = 0.5j  # magnetic dipole along x
    I, E = O.intensity(dirs, Fm)
    assert np.allclose(I, z0 ** 2 * 0.25 * (1 - dirs[:, 0] ** 2), rtol=1e-13, atol=1e-12)
    assert np.abs(np.einsum("cd,dc->d", E, dirs)).max() < 1e-9


# -------------------------------------------------------------- octree -----
@pytest.mark.parametrize("depth", [0, 1, 2, 4])
def test_octree_partition_invariants_and_sort_restatement(depth):
    rng = np.random.default_rng(depth + 11)
    n = 5000
    x, y, z = [rng.uniform(-1.2, 1.2, n).astype(np.float32) * np.float32(1e-6) for _ in range(3)]
    x[:40] = 0.0  # on the root split plane
    box = np.array([-1e-6, -1e-6, -1e-6, 1e-6, 1e-6, 1e-6], np.float32)  # some points outside
    perm, nb, ne, nbox = O.octree([x, y, z], box, depth)
    assert sorted(perm.tolist()) == list(range(n))
    # independent restatement: key from float split planes, stable sort by key
    key = np.zeros(n, np.int64)
    lo = np.tile(box[:3], (n, 1)).astype(np.float32)
    hi = np.tile(box[3:], (n, 1)).astype(np.float32)
    P = np.stack([x, y, z], 1)
    for _ in range(depth):
        c = (np.float32(0.5) * (lo + hi)).astype(np.float32)
        ge = P >= c
        dig = 4 * (~ge[:, 2]) + 2 * (~ge[:, 1]) + ge[:, 0]
        key = key * 8 + dig
        lo = np.where(ge, c, lo)
        hi = np.where(ge, hi, c)
    order = np.argsort(key, kind="stable")
    np.testing.assert_array_equal(perm, order)
    # ranges: each level partitions [0, n) in node order, children concatenate
    off = 0
    for lvl in range(depth + 1):
        cnt = 8 ** lvl
        b, e = nb[off:off + cnt], ne[off:off + cnt]
        assert b[0] == 0 and e[-1] == n and np.all(b[1:] == e[:-1]) and np.all(e >= b)
        if lvl < depth:
            cb, ce = nb[off + cnt:off + 9 * cnt], ne[off + cnt:off + 9 * cnt]
            assert np.all(cb.reshape(-1, 8)[:, 0] == b) and np.all(ce.reshape(-1, 8)[:, -1] == e)
        off += cnt
    # points inside the root box lie in their leaf box (half-open)
    lo_leaf = 8 ** depth
    loff = (8 ** depth - 1) // 7
    for j in range(lo_leaf):
        sl = perm[nb[loff + j]:ne[loff + j]]
        bx = nbox[loff + j]
        inside = np.all((P[sl] >= box[:3]) & (P[sl] < box[3:]), 1)
        q = P[sl][inside]
        assert np.all((q >= bx[:3]) & (q <= bx[3:]))


# ------------------------------------------------------ level sampling -----
def test_reference_level_counts_appendix_b():
    # sphere r = 10 um, rho = 1.3 (sphere.cu:346-350): extent 2 r rho
    assert [O.ref_level_count(i, np.float32(2 * 10e-6 * 1.3)) for i in range(1, 7)] == \
        [240, 131, 74, 44, 27, 18]
    # fiber index 1 (treePO.cu:44-51): theta extent period*rho, phi extent 2 r rho
    assert [O.ref_level_count(i, np.float32(50e-6 * 1.2)) for i in range(1, 7)] == \
        [518, 273, 148, 83, 48, 30]
    assert [O.ref_level_count(i, np.float32(2 * 40e-6 * 1.2)) for i in range(1, 7)] == \
        [809, 421, 223, 122, 69, 41]


# ------------------------------------------------------- interpolation -----
@pytest.mark.parametrize("nc,np_", [(7, 16), (8, 15), (12, 12), (16, 9), (9, 4)])
def test_trig_interp_vs_numpy_fft(nc, np_):
    rng = np.random.default_rng(nc * 31 + np_)
    v = rng.standard_normal(nc) + 1j * rng.standard_normal(nc)
    off_c, off_p = 0.3, -0.2
    out = O.trig_interp(v, off_c, np_, off_p)
    # independent: numpy FFT coefficients, evaluated explicitly
    X = np.fft.fft(v) / nc
    f = np.fft.fftfreq(nc, 1 / nc)
    t = off_p + 2 * np.pi * np.arange(np_) / np_ - off_c
    ref = np.zeros(np_, complex)
    for m in range(nc):
        if nc % 2 == 0 and m == nc // 2:
            ref += 0.5 * X[m] * (np.exp(1j * (nc // 2) * t) + np.exp(-1j * (nc // 2) * t))
        else:
            ref += X[m] * np.exp(1j * f[m] * t)
    assert np.abs(out - ref).max() < 1e-12 * max(1, np.abs(ref).max())
    if np_ == nc and off_c == off_p:
        assert np.abs(out - v).max() < 1e-12


def test_trig_interp_exact_on_band_limited():
    nc, np_ = 11, 40
    coef = {-4: 0.3 + 1j, -1: 2.0, 0: -0.5j, 2: 1.5, 5: 0.25 - 0.5j}
    f = lambda t: sum(a * np.exp(1j * k * t) for k, a in coef.items())
    off_c = 0.1
    v = f(off_c + 2 * np.pi * np.arange(nc) / nc)
    out = O.trig_interp(v, off_c, np_, 0.0)
    assert np.abs(out - f(2 * np.pi * np.arange(np_) / np_)).max() < 1e-12


@pytest.mark.parametrize("nc, nup", [(9, 24), (12, 30), (15, 15)])
def test_trig_interp_up_then_down_identity(nc, nup):
    """Resampling an arbitrary nc-sample line onto a finer grid (nup >= nc) and
    back recovers it (the finer grid holds every frequency of the coarse one),
    including a shifted origin; the round trip through a coarser grid does not
    (it drops frequencies), so the identity is not trivially met."""
    rng = np.random.default_rng(nc)
    v = rng.standard_normal(nc) + 1j * rng.standard_normal(nc)
    off_c, off_u = 0.37, -0.21
    up = O.trig_interp(v, off_c, nup, off_u)
    back = O.trig_interp(up, off_u, nc, off_c)
    assert np.abs(back - v).max() < 1e-12 * np.abs(v).max()
    if nc > 4:
        down = O.trig_interp(v, off_c, nc - 3, off_u)
        again = O.trig_interp(down, off_u, nc, off_c)
        assert np.abs(again - v).max() > 1e-3 * np.abs(v).max()


def test_sphere_interp_point_sources():
    """Pattern of sources within radius R is band-limited (degree ~ kR): the 2 pi
    theta extension + phi interpolation must reproduce it on a new grid."""
    rng = np.random.default_rng(2)
    src = rng.uniform(-1, 1, (5, 3)) * 0.6
    amp = rng.standard_normal(5) + 1j * rng.standard_normal(5)

    def pat(t, p):
        T, P = np.meshgrid(t, p, indexing="ij")
        s = np.stack([np.cos(T) * np.cos(P), np.cos(T) * np.sin(P), np.sin(T)], -1)
        return np.einsum("s,ijs->ij", amp, np.exp(1j * s @ src.T))

    ntc, npc = 16, 30
    tc = -np.pi / 2 + np.pi / (2 * ntc) + np.pi * np.arange(ntc) / ntc
    child = pat(tc, 2 * np.pi * np.arange(npc) / npc)
    for ntp, npp, offp in [(25, 48, -np.pi / 2 + np.pi / 50), (12, 20, -np.pi / 2 + 1e-10)]:
        out = O.sphere_interp(child, ntp, npp, offp)
        ref = pat(offp + np.pi * np.arange(ntp) / ntp, 2 * np.pi * np.arange(npp) / npp)
        assert np.abs(out - ref).max() < 1e-7 * np.abs(ref).max()


# ------------------------------------------------------ reference fixture --
def test_ellipsoid_fixture_properties():
    import os

    g = os.path.join(os.path.dirname(__file__), "golden")
    n = 30322
    p = np.fromfile(os.path.join(g, "ellipsoid_3_3_2_float.binary"), np.float32).reshape(3, n)
    q = np.fromfile(os.path.join(g, "ellipsoid_3_3_2_normal_float.binary"), np.float32).reshape(3, n)
    a = np.fromfile(os.path.join(g, "ellipsoid_3_3_2_area_float.binary"), np.float32)
    assert abs(a.sum() / 8.8931e-11 - 1) < 1e-4
    assert np.abs(np.linalg.norm(q, axis=0) - 1).max() < 1e-6
    # semi-axes 3, 3, 2 um
    assert abs(p[0].max() - 3e-6) < 1e-8 and abs(p[2].max() - 2e-6) < 1e-8
    # under +z incidence exactly two samples take the normal-incidence branch
    assert int(np.sum(np.abs(np.abs(q[2].astype(np.float64)) - 1) < 1e-10)) == 2


def test_normal_slightly_longer_than_one_stays_finite():
    """A normal of length 1 + 1e-7 along the incidence direction (|n.w| > 1 by
    rounding) takes the normal-incidence branch: finite currents equal to
    those of the unit normal scaled by the tiny length change."""
    import oracle as orc

    pts = [np.array([0.0]), np.array([0.0]), np.array([0.0]),
           np.array([0.0]), np.array([0.0]), np.array([-1.0 - 1e-7]), np.array([1e-12])]
    unit = [a.copy() for a in pts]
    unit[5] = np.array([-1.0])
    inc = (np.pi / 2, 0.0, 0.5e-6, "M", 1.55, 0.0)
    c = orc.currents(pts, inc)
    cu = orc.currents(unit, inc)
    assert np.all(np.isfinite(c))
    assert np.allclose(c, cu, rtol=1e-6, atol=1e-6 * np.abs(cu).max())
