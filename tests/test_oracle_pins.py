"""Pins of the FP64 CPU oracle against independent answers.

The reference ships no tests or golden vectors for the photon path, so each
part of the oracle is pinned here by something it cannot share a mistake
with: published known-answer vectors, libm, closed forms, invariants and
brute force on small inputs.
"""
import ctypes as C
import math

import numpy as np
import pytest

from kratos_polrad_amd import _abi, configs
from kratos_polrad_amd.engine import BandSpec, RunSpec
from kratos_polrad_amd.mesh import Mesh


# --------------------------------------------------------------------------
# Philox4x32-10: Random123 known-answer vectors (Salmon et al. 2011,
# kat_vectors "philox4x32 10" lines).
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF),
     (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(oracle_lib, ctr, key, want):
    assert tuple(oracle_lib.philox(ctr, *key)) == want


def test_philox_counter_sensitivity(oracle_lib):
    # distinct packet ids / events give uncorrelated words (avalanche check)
    a = oracle_lib.philox((7, 0, 1, 0), 42, 0)
    b = oracle_lib.philox((7, 0, 2, 0), 42, 0)
    c = oracle_lib.philox((7, 0, 1, 0), 42, 1)
    bits = lambda x, y: sum(bin(p ^ q).count("1") for p, q in zip(x, y))
    assert 40 < bits(a, b) < 88 and 40 < bits(a, c) < 88


# --------------------------------------------------------------------------
# Deterministic transcendentals against libm
def _ulp_err(x, ref):
    return abs(x - ref) / max(math.ulp(ref), 1e-300)


def test_log_matches_libm(oracle_lib):
    L = oracle_lib.lib()
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(1e-10, 1, 20000), (np.arange(1, 2000) + 0.5) * 2.0 ** -32,
                         [1.0, 0.5, 2.0 ** -32, 0.7071067811865476, 0.7071067811865475]])
    worst = max(_ulp_err(L.or_log(float(x)), math.log(float(x))) for x in xs)
    assert worst <= 4.0


def test_exp_matches_libm(oracle_lib):
    L = oracle_lib.lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([-rng.uniform(0, 60, 20000), -rng.uniform(0, 1e-3, 1000), [0.0, -700.0]])
    worst = max(_ulp_err(L.or_exp(float(x)), math.exp(float(x))) for x in xs)
    assert worst <= 4.0
    assert L.or_exp(-800.0) == 0.0


def test_sincos2pi_matches_libm(oracle_lib):
    """Against x87 extended precision (2 pi u rounds in double, so a double
    libm reference is itself ~1e-15 off)."""
    L = oracle_lib.lib()
    s, c = C.c_double(), C.c_double()
    rng = np.random.default_rng(2)
    two_pi = 2 * np.longdouble("3.14159265358979323846264338327950288")
    for u in list(rng.uniform(0, 1, 20000)) + [0.0, 0.125, 0.25, 0.5, 0.75, 0.999999]:
        L.or_sincos2pi(float(u), C.byref(s), C.byref(c))
        a = two_pi * np.longdouble(u)
        assert abs(np.longdouble(s.value) - np.sin(a)) < 3e-16
        assert abs(np.longdouble(c.value) - np.cos(a)) < 3e-16


# --------------------------------------------------------------------------
# Scattering-angle sampling
def test_thomson_mu_inverts_cdf(oracle_lib):
    L = oracle_lib.lib()
    xi = (np.arange(200000) + 0.5) / 200000
    mu = np.array([L.or_thomson_mu(float(x)) for x in xi])
    # CDF (3/8)[(mu+1) + (mu^3+1)/3] equals xi
    cdf = 0.375 * ((mu + 1) + (mu ** 3 + 1) / 3)
    assert np.max(np.abs(cdf - xi)) < 1e-13
    # moments of (3/8)(1+mu^2): <mu> = 0, <mu^2> = 2/5
    assert abs(mu.mean()) < 1e-6
    assert abs((mu ** 2).mean() - 0.4) < 1e-6


@pytest.mark.parametrize("g", [-0.5, 0.3, 0.6, 0.9])
def test_hg_mu_mean_is_g(oracle_lib, g):
    L = oracle_lib.lib()
    xi = (np.arange(400000) + 0.5) / 400000
    mu = np.array([L.or_hg_mu(g, float(x)) for x in xi])
    assert abs(mu.mean() - g) < 2e-4
    # CDF of HG: (1-g^2)/(2g) [ (1+g^2-2g mu)^-1/2 - (1+g)^-1 ]
    cdf = (1 - g * g) / (2 * g) * ((1 + g * g - 2 * g * mu) ** -0.5 - 1 / (1 + g))
    assert np.max(np.abs(cdf - xi)) < 1e-9


# --------------------------------------------------------------------------
# Stokes algebra: sign conventions pinned by physics
def _pol_vector(d, e1, S):
    """Physical polarization direction (unit 3-vector) and degree."""
    e2 = np.cross(d, e1)
    chi = 0.5 * math.atan2(S[2], S[1])
    return math.cos(chi) * e1 + math.sin(chi) * e2, math.hypot(S[1], S[2]) / S[0]


def test_thomson_90deg_fully_polarized_perpendicular(oracle_lib):
    d = np.array([0.0, 0.0, 1.0])
    e1 = np.array([1.0, 0.0, 0.0])
    for phi in (0.0, 0.1, 0.3, 0.77):
        dn, en, S = oracle_lib.scatter_stokes(_abi.KP_THOMSON, 1.0, 0.0, phi, d, e1,
                                              (1.0, 0.0, 0.0))
        assert abs(np.dot(dn, d)) < 1e-15  # 90 degrees
        pv, p = _pol_vector(dn, en, S)
        assert abs(p - 1.0) < 1e-12
        # perpendicular to the scattering plane spanned by d and dn
        assert abs(np.dot(pv, d)) < 1e-12 and abs(np.dot(pv, dn)) < 1e-12
        assert S[0] == 1.0  # unpolarized input keeps I under weighting


def test_hg_rayleigh_degree_is_pmax_at_90(oracle_lib):
    d, e1 = np.array([0.0, 1.0, 0.0]), np.array([0.0, 0.0, 1.0])
    _, en, S = oracle_lib.scatter_stokes(_abi.KP_HG_RAYLEIGH, 0.5, 0.0, 0.2, d, e1,
                                         (1.0, 0.0, 0.0))
    assert abs(math.hypot(S[1], S[2]) / S[0] - 0.5) < 1e-12


def test_scatter_is_frame_independent(oracle_lib):
    """Rotating the input reference axis (and the Stokes vector with it) must not
    change the outgoing physical state: direction, I, and polarization vector."""
    rng = np.random.default_rng(5)
    for _ in range(50):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        e1 = np.cross(d, rng.normal(size=3))
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(d, e1)
        q, u = rng.uniform(-0.6, 0.6, 2)
        S = np.array([1.0, q, u])
        mu, phi = rng.uniform(-1, 1), rng.uniform(0, 1)
        dn, en, Sn = oracle_lib.scatter_stokes(_abi.KP_THOMSON, 1.0, mu, phi, d, e1, S)
        psi = rng.uniform(0, 2 * math.pi)
        e1r = math.cos(psi) * e1 + math.sin(psi) * e2
        c2, s2 = math.cos(2 * psi), math.sin(2 * psi)
        Sr = np.array([1.0, q * c2 + u * s2, u * c2 - q * s2])
        # same physical scattering: azimuth measured from the rotated axis
        phir = (phi - psi / (2 * math.pi)) % 1.0
        dn2, en2, Sn2 = oracle_lib.scatter_stokes(_abi.KP_THOMSON, 1.0, mu, phir, d, e1r, Sr)
        assert np.allclose(dn, dn2, atol=1e-10)
        assert abs(Sn[0] - Sn2[0]) < 1e-10
        p1, deg1 = _pol_vector(dn, en, Sn)
        p2, deg2 = _pol_vector(dn2, en2, Sn2)
        assert abs(deg1 - deg2) < 1e-9
        if deg1 > 1e-6:
            assert abs(abs(np.dot(p1, p2)) - 1.0) < 1e-8


def test_peel_matches_scatter_toward_observer(oracle_lib):
    """Peeling toward n_o equals scattering exactly into n_o, expressed in
    the image frame, times the phase function Z11."""
    rng = np.random.default_rng(9)
    no = np.array([1.0, 0.0, 0.0])
    ev = np.array([0.0, 0.0, 1.0])
    for _ in range(30):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        e1 = np.cross(d, rng.normal(size=3))
        e1 /= np.linalg.norm(e1)
        S = np.array([1.0, *rng.uniform(-0.5, 0.5, 2)])
        Si = oracle_lib.peel_stokes(_abi.KP_THOMSON, 0.0, 1.0, d, e1, S, no, ev)
        mu = float(np.dot(d, no))
        e2 = np.cross(d, e1)
        p = no - mu * d
        phi = math.atan2(np.dot(p, e2), np.dot(p, e1)) / (2 * math.pi) % 1.0
        dn, en, Sn = oracle_lib.scatter_stokes(_abi.KP_THOMSON, 1.0, mu, phi, d, e1, S)
        assert np.allclose(dn, no, atol=1e-9)
        z11 = 0.75 * (1 + mu * mu)
        assert abs(Si[0] - z11 * Sn[0]) < 1e-9
        # same polarization vector in 3D
        pv1, deg1 = _pol_vector(dn, en, Sn)
        pv2, deg2 = _pol_vector(no, ev, Si)
        assert abs(deg1 - deg2) < 1e-9
        if deg1 > 1e-6:
            assert abs(abs(np.dot(pv1, pv2)) - 1.0) < 1e-8


# --------------------------------------------------------------------------
# March: reference semantics and geometric invariants
def _amr_mesh(n_cell=4):
    m = Mesh(root_blocks=(3, 2, 2), n_cell=n_cell, x_min=(0, 0, 0), x_max=(1.5, 1.0, 1.0))
    m.refine_region(1, (0.4, 0.3, 0.3), (0.9, 0.7, 0.6))
    m.refine_region(2, (0.55, 0.45, 0.4), (0.7, 0.55, 0.5))
    return m.finalize()


def _brute_locate(m, x):
    hi = m.xf0 + m.n_cell * m.dx0
    inside = np.all((x >= m.xf0) & (x < hi), axis=1)
    b = np.nonzero(inside)[0]
    assert len(b) == 1, (x, b)
    b = b[0]
    i = np.floor((x - m.xf0[b]) / m.dx0[b]).astype(int)
    return b, i[0] + m.n_cell * (i[1] + m.n_cell * i[2])


def _chord(x, d, lo, hi):
    t = np.inf
    for a in range(3):
        if d[a] > 0:
            t = min(t, (hi[a] - x[a]) / d[a])
        elif d[a] < 0:
            t = min(t, (lo[a] - x[a]) / d[a])
    return t


def test_march_segments_match_point_location(oracle_lib):
    m = _amr_mesh()
    rng = np.random.default_rng(11)
    for _ in range(300):
        x = rng.uniform([0.01, 0.01, 0.01], [1.49, 0.99, 0.99])
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        ib, cell, ln, xe, st = oracle_lib.trace(m, x, d)
        assert st == 1
        assert abs(ln.sum() - _chord(x, d, m.x_min, m.x_max)) < 1e-12
        t = 0.0
        for b, c, l in zip(ib, cell, ln):
            if l > 1e-9:
                bb, cc = _brute_locate(m, x + (t + 0.5 * l) * d)
                assert (bb, cc) == (b, c)
            t += l


def test_tie_goes_to_highest_axis(oracle_lib):
    """photon.h:38 (`dl <= dl_min`): through a cell corner the highest axis is
    crossed first, then a zero-length step crosses the other."""
    m = Mesh(root_blocks=(1, 1, 1), n_cell=4).finalize()
    # from the centre of cell (0,0,0) along (1,1,0)/sqrt2: hits the x-y edge
    x = (0.125, 0.125, 0.125)
    d = (math.sqrt(0.5), math.sqrt(0.5), 0.0)
    ib, cell, ln, xe, st = oracle_lib.trace(m, x, d)
    assert cell[0] == 0
    assert cell[1] == 0 + 4 * 1  # y crossed first (axis 1 > axis 0)
    assert ln[1] == 0.0          # then the zero-length x crossing
    assert cell[2] == 1 + 4 * 1


def test_periodic_wrap(oracle_lib):
    m = Mesh(root_blocks=(2, 1, 1), n_cell=4, periodic=(1, 0, 0)).finalize()
    x = (0.9, 0.5, 0.5)
    d = (math.cos(0.3), 0.0, math.sin(0.3))
    ib, cell, ln, xe, st = oracle_lib.trace(m, x, d)
    assert st == 1
    # leaves through z=1 after wrapping in x; total length = chord in z
    assert abs(ln.sum() - 0.5 / math.sin(0.3)) < 1e-12
    assert 0.0 <= xe[0] < 1.0


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_periodic_wrap_both_directions_no_zero_steps(oracle_lib, sign):
    """A ray crossing the periodic faces in either direction visits every
    cell on its line exactly as in the unwrapped domain: step lengths sum to
    the travelled distance, no zero-length steps, cells visited match point
    location at each step midpoint (lower-face crossings included)."""
    m = Mesh(root_blocks=(2, 2, 1), n_cell=4, periodic=(1, 1, 0))
    m.refine(1, [[0, 0, 0], [3, 3, 1]])
    m.finalize()
    x = (0.31, 0.62, 0.13)
    d = np.array([sign * 0.8, sign * -0.5, 0.33])
    d /= np.linalg.norm(d)
    ib, cell, ln, xe, st = oracle_lib.trace(m, x, tuple(d))
    assert st == 1  # leaves through z = 1
    assert np.all(ln > 0)
    assert abs(ln.sum() - (1.0 - 0.13) / d[2]) < 1e-12
    # every step's midpoint lies in the cell the march assigned
    pos = np.array(x, dtype=np.float64)
    for k in range(len(ln)):
        mid = pos + 0.5 * ln[k] * d
        w = mid.copy()
        for a in range(2):
            w[a] %= 1.0
        b, i = m.locate(w)
        assert (b, i[0] + 4 * (i[1] + 4 * i[2])) == (ib[k], cell[k]), k
        pos = pos + ln[k] * d


# --------------------------------------------------------------------------
# Transport: closed forms
def _uniform_media(m, kappa):
    n = m.n_cells
    return (np.full(n, kappa, np.float32), np.ones(n, np.float32), np.zeros(n, np.float32))


def test_optically_thin_tally_closed_form(oracle_lib):
    """kappa = 0: the path-length tally (photon.h:73-83) has expectation
    E[J_c] = N w / (4 pi V_c) * integral_c dV / r^2 (cubature per cell)."""
    m = Mesh(root_blocks=(2, 2, 2), n_cell=4).finalize()
    k, a, g = _uniform_media(m, 0.0)
    src = (0.37, 0.52, 0.61)
    N, nb = 40000, 8
    Js = []
    for b in range(nb):
        run = RunSpec(bands=[BandSpec(n_packets=N)], packet_offset=b * N, incl_deg=())
        Js.append(oracle_lib.forward(m, k, a, g, src, run, image=False)["J"])
    Js = np.array(Js)
    mean, err = Js.mean(0), Js.std(0, ddof=1) / np.sqrt(nb)
    # cubature of 1/r^2 over each cell (12^3 midpoints)
    q = (np.arange(12) + 0.5) / 12
    qq = np.stack(np.meshgrid(q, q, q, indexing="ij"), -1).reshape(-1, 3)
    xc = m.cell_centers()
    h = np.repeat(m.dx0, m.n3, axis=0)
    exp = np.empty(m.n_cells)
    for c in range(m.n_cells):
        pts = xc[c] - 0.5 * h[c] + qq * h[c]
        exp[c] = N / (4 * np.pi) * np.mean(1.0 / np.sum((pts - src) ** 2, axis=1))
    far = np.linalg.norm(xc - src, axis=1) > 0.2
    z = (mean[far] - exp[far]) / np.maximum(err[far], 1e-12)
    assert np.all(np.abs(z) < 5.0)
    assert abs(mean[far].sum() / exp[far].sum() - 1) < 3e-3


def test_free_flight_escape_probability(oracle_lib):
    """Uniform absorber (albedo 0): a packet escapes iff its first free flight
    exceeds the distance to the boundary, P = exp(-kappa d).  The emission
    direction is recomputed from the published Philox words."""
    m = Mesh(root_blocks=(1, 1, 1), n_cell=8).finalize()
    kap = 3.0
    k, a, g = _uniform_media(m, kap)
    src = np.array([0.3, 0.55, 0.6])
    N = 20000
    run = RunSpec(bands=[BandSpec(n_packets=N, albedo=0.0)], seed=77, incl_deg=(),
                  rr_weight=0.5)
    out = oracle_lib.forward(m, k, a, g, tuple(src), run, image=False, end_state=True)
    es = out["end_state"]
    # survivors of roulette after absorption carry zero weight; a packet that
    # never interacted escapes with its emitted weight
    escaped = (es[:, _abi.KP_ES_STATUS] == _abi.KP_END_ESCAPED) & (es[:, _abi.KP_ES_I] > 0)
    assert np.all(es[escaped, _abi.KP_ES_I] == 1.0)
    assert np.all(es[escaped, _abi.KP_ES_EVENTS] == 0)
    P = np.empty(N)
    L = oracle_lib.lib()
    s, c = C.c_double(), C.c_double()
    for p in range(N):
        w = oracle_lib.philox((p, 0, 0, 0), 77, 0)
        u = [(x + 0.5) * 2.0 ** -32 for x in w]
        mu = 2 * u[0] - 1
        L.or_sincos2pi(u[1], C.byref(s), C.byref(c))
        st = math.sqrt(max(0.0, 1 - mu * mu))
        d = np.array([st * c.value, st * s.value, mu])
        P[p] = math.exp(-kap * _chord(src, d, m.x_min, m.x_max))
    expect, sd = P.sum(), math.sqrt((P * (1 - P)).sum())
    assert abs(escaped.sum() - expect) < 4 * sd


def test_energy_conservation_albedo_one(oracle_lib):
    w = configs.config1(n_packets=100000)
    out = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, w.run, end_state=True,
                             image=False)
    es = out["end_state"]
    esc = es[:, _abi.KP_ES_STATUS] == _abi.KP_END_ESCAPED
    # Russian roulette is unbiased: escaped weight / emitted weight -> 1
    r = es[esc, _abi.KP_ES_I].sum() / 100000
    assert abs(r - 1) < 3e-3
    assert out["counters"][_abi.KP_C_LOST] == 0


def test_pure_absorber_peel_is_exact(oracle_lib):
    """Albedo 0: the image is only the emission peel-off w e^{-tau}/(4 pi), with
    tau the optical depth of the streamed ray to the edge."""
    w = configs.config1(n_packets=5000)
    w.run.bands[0].albedo = 0.0
    w.run.incl_deg = (30.0, 45.0, 90.0)
    w.run.image_half = 0.9
    out = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, w.run)
    L = oracle_lib.lib()
    kap = w.kappa.reshape(w.mesh.n_blocks, -1)
    s, c = C.c_double(), C.c_double()
    for o, inc in enumerate(w.run.incl_deg):
        L.or_sincos2pi(inc / 360.0, C.byref(s), C.byref(c))
        ib, cell, ln, xe, st = oracle_lib.trace(w.mesh, w.source, (s.value, 0.0, c.value))
        tau = float(sum(kap[b, cc] * l for b, cc, l in zip(ib, cell, ln)))
        exact = 5000 * math.exp(-tau) / (4 * math.pi)
        assert abs(out["image"][0, o, ..., 0].sum() / exact - 1) < 1e-12
        assert np.all(out["image"][0, o, ..., 1:] == 0)


def test_thomson_sphere_tangential_polarization(oracle_lib):
    """Centred source in a scattering sphere: polarization is tangential
    (perpendicular to the projected radius) and Q, U sum to ~0."""
    w = configs.config1(n_packets=200000)
    out = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, w.run)
    img = out["image"][0, 0]
    I, Q, U = img[..., 0], img[..., 1], img[..., 2]
    up = (slice(90, 110), slice(62, 66))      # above centre on the vertical axis
    right = (slice(62, 66), slice(90, 110))   # right of centre on the horizontal axis
    assert Q[up].sum() / I[up].sum() < -0.2
    assert Q[right].sum() / I[right].sum() > 0.2
    assert abs(U[up].sum()) < 0.05 * I[up].sum()
    d1 = (slice(85, 100), slice(85, 100))
    d2 = (slice(85, 100), slice(28, 43))
    assert U[d1].sum() * U[d2].sum() < 0  # antisymmetric on the diagonals
    assert abs(Q.sum()) < 0.01 * I.sum() and abs(U.sum()) < 0.01 * I.sum()
