"""Spherical meshes and the spherical march: pins of the CPU restatement.

The reference's spherical photon (radiation/photon_sph.h) has no tests or
golden outputs, so the restatement in oracle/kp_sph_oracle.c is pinned by
what the geometry fixes:
  - cell volumes of meshgen/spherical.h:31-42 sum to the shell / wedge volume;
    the r axis of cylindrical.h:34-43 has a constant width ratio;
  - the march's own point location agrees with kp_mesh_locate of the
    (r, theta, phi) coordinates from libm (acos, atan2);
  - a radial ray crosses only r faces, each step one cell width less the
    1.001 overshoot of the step before;
  - on random rays every tallied segment's midpoint lies in the cell it was
    tallied to (independent libm point location), and 1.001 * sum(l) equals
    the chord to the domain boundary plus 0.001 of the last step;
  - a mirror image in z visits the mirrored theta cells with equal lengths
    (the reference's theta = pi/2 - z/r approximation for z/r < 0.01 would
    break this; the restatement uses the exact polar angle);
  - per radial shell, the tallied sum of J*V equals the exact length of the
    packets' rays inside the shell (ray/sphere intersections from the
    directions in the end states), divided by the overshoot factor.
"""
import numpy as np
import pytest

import oracle
from kratos_polrad_amd import _abi, configs
from kratos_polrad_amd.engine import BandSpec, RunSpec
from kratos_polrad_amd.mesh import Mesh

TWO_PI = 2.0 * np.pi


def _mesh(refine=True, r_ratio=-1.0, root=(4, 2, 4), n_cell=8, theta=(0.0, np.pi),
          phi=(0.0, TWO_PI)):
    ref = ()
    if refine:
        ref = ((1, (0.2, 1.0, 1.0), (0.4, 1.5, 2.0)),
               (2, (0.25, 1.2, 1.3), (0.3, 1.3, 1.5)))
    return configs.sph_mesh(root_blocks=root, n_cell=n_cell, r_lim=(0.1, 1.0), r_ratio=r_ratio,
                            theta=theta, phi=phi, refine=ref)


def _sph_coords(x):
    r = float(np.sqrt(x @ x))
    th = float(np.arccos(np.clip(x[2] / r, -1.0, 1.0)))
    ph = float(np.arctan2(x[1], x[0]) % TWO_PI)
    return np.array([r, th, ph])


def _random_points(rng, n, r_lo=0.1, r_hi=1.0):
    out = []
    while len(out) < n:
        x = rng.uniform(-r_hi, r_hi, 3)
        if r_lo < np.sqrt(x @ x) < r_hi:
            out.append(x)
    return np.array(out)


def _near_face(m, ib, rtp, tol=1e-9):
    f = m.xf[ib]
    return any(np.min(np.abs(f[a] - rtp[a])) < tol * max(1.0, abs(rtp[a])) for a in range(3))


def test_volumes_sum_to_shell_and_wedge():
    m = _mesh()
    v = m.cell_volumes()
    assert np.all(v > 0)
    assert v.sum() == pytest.approx(4.0 / 3.0 * np.pi * (1.0 - 0.1 ** 3), rel=1e-12)
    w = _mesh(refine=False, theta=(np.pi / 4, 3 * np.pi / 4), phi=(0.0, np.pi / 2))
    expect = (1.0 - 0.1 ** 3) / 3.0 * (np.cos(np.pi / 4) - np.cos(3 * np.pi / 4)) * (np.pi / 2)
    assert w.cell_volumes().sum() == pytest.approx(expect, rel=1e-12)
    assert w.periodic == (0, 0, 0) and m.periodic == (0, 0, 1)


def test_log_r_axis_and_face_tables():
    m = _mesh(refine=False, root=(4, 2, 4), n_cell=8)
    r = np.unique(m.xf[:, 0, :].ravel())
    assert r[0] == pytest.approx(0.1) and r[-1] == pytest.approx(1.0)
    q = r[1:] / r[:-1]
    assert np.allclose(q, (1.0 / 0.1) ** (1.0 / 32), rtol=1e-12)
    # the table's trigonometry is that of the face coordinates
    n, mm = m.n_cell, m.n_cell + 1
    T = m.sph
    assert np.array_equal(T[:, :mm], m.xf[:, 0, :])
    assert np.allclose(T[:, mm:2 * mm], np.cos(m.xf[:, 1, :]), atol=1e-15)
    assert np.allclose(T[:, 2 * mm:3 * mm], np.sin(m.xf[:, 2, :] % TWO_PI), atol=1e-15)
    assert np.allclose(T[:, 3 * mm:4 * mm], np.cos(m.xf[:, 2, :] % TWO_PI), atol=1e-15)
    # r ratio 1: uniform r axis with a positive dx0
    u = _mesh(refine=False, r_ratio=1.0)
    assert np.all(u.dx0[:, 0] > 0) and np.all(m.dx0[:, 0] == -1.0)
    # (a uniform axis is xf0 + i*dx0 per block: block edges agree to rounding)
    assert np.allclose(np.diff(np.unique(np.round(u.xf[:, 0, :].ravel(), 12))), 0.9 / 32,
                       rtol=1e-9)


def test_set_coord_validation():
    with pytest.raises(_abi.KPError):
        Mesh(root_blocks=(2, 2, 2), n_cell=8, x_min=(0.0, 0.0, 0.0),
             x_max=(1.0, np.pi, TWO_PI), coord="spherical")  # r_min = 0
    with pytest.raises(_abi.KPError):
        Mesh(root_blocks=(2, 2, 1), n_cell=2, x_min=(0.1, 0.0, 0.0),
             x_max=(1.0, np.pi, TWO_PI), periodic=(0, 0, 1), coord="spherical")  # phi cells = pi
    with pytest.raises(_abi.KPError):
        Mesh(root_blocks=(2, 2, 2), n_cell=8, x_min=(0.1, 0.0, 0.0),
             x_max=(1.0, np.pi, 3.0), periodic=(0, 0, 1), coord="spherical")  # periodic != 2pi


def test_march_locate_matches_mesh_locate():
    m = _mesh()
    rng = np.random.default_rng(3)
    n_checked = 0
    for x in _random_points(rng, 400):
        rtp = _sph_coords(x)
        ib, i = m.locate(rtp)
        if ib < 0 or _near_face(m, ib, rtp):
            continue
        assert oracle.sph_locate(m, x) == (ib, i)
        n_checked += 1
    assert n_checked > 350


def test_radial_ray_steps_are_cell_widths():
    m = _mesh(refine=False)
    d = np.array([0.31, -0.47, 0.22])
    d /= np.linalg.norm(d)
    x = 0.1 * (1 + 1e-9) * d
    ib, cell, ln, xe, st = oracle.sph_trace(m, x, d)
    assert st == 1  # escaped through r = 1
    n = m.n_cell
    r = np.unique(m.xf[:, 0, :].ravel())
    assert len(ln) == len(r) - 1
    i1 = (cell // n) % n
    i2 = cell // (n * n)
    assert len(set(zip(ib[np.diff(np.r_[ib, -1]) != 0].tolist()))) >= 4
    assert np.all(i1 == i1[0]) and np.all(i2 == i2[0])  # no theta / phi crossing
    widths = np.diff(r)
    expect = widths.copy()
    expect[0] = r[1] - 0.1 * (1 + 1e-9)
    expect[1:] -= 0.001 * ln[:-1]
    assert np.allclose(ln, expect, rtol=1e-9, atol=1e-15)
    # inward along the same line: leaves through the inner sphere
    ib2, cell2, ln2, xe2, st2 = oracle.sph_trace(m, 0.999 * d, -d)
    assert st2 == 1 and np.linalg.norm(xe2) < 0.1


def _chord(x, d, r_in, r_out):
    b = x @ d
    c_in = x @ x - r_in * r_in
    det = b * b - c_in
    if det > 0 and -b - np.sqrt(det) > 0:
        return -b - np.sqrt(det)
    return -b + np.sqrt(b * b - (x @ x - r_out * r_out))


@pytest.mark.parametrize("refine", [False, True])
def test_random_rays_midpoints_and_chords(refine):
    m = _mesh(refine=refine)
    n = m.n_cell
    rng = np.random.default_rng(11 + refine)
    checked = 0
    for x0 in _random_points(rng, 150):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        ib, cell, ln, xe, st = oracle.sph_trace(m, x0, d)
        assert st == 1 and len(ln) > 0 and np.all(ln > 0)
        ch = _chord(x0, d, 0.1, 1.0)
        assert 1.001 * ln.sum() - ch == pytest.approx(0.001 * ln[-1], rel=1e-6, abs=1e-13)
        x = x0.copy()
        for k in range(len(ln)):
            mid = _sph_coords(x + 0.5 * ln[k] * d)
            jb, j = m.locate(mid)
            if not _near_face(m, jb, mid, tol=1e-7):
                assert (jb, j[0] + n * (j[1] + n * j[2])) == (ib[k], cell[k]), (k, mid)
                checked += 1
            x = x + (ln[k] * d) * 1.001
        assert np.allclose(x, xe, rtol=0, atol=1e-15)
    assert checked > 1000


def test_mirror_symmetry_in_z():
    m = _mesh(refine=False, root=(4, 2, 4))
    n_th = 2 * m.n_cell
    rng = np.random.default_rng(5)
    n = m.n_cell
    for x0 in _random_points(rng, 60):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        flip = np.array([1.0, 1.0, -1.0])
        a = oracle.sph_trace(m, x0, d)
        b = oracle.sph_trace(m, x0 * flip, d * flip)
        assert len(a[2]) == len(b[2])
        assert np.allclose(a[2], b[2], rtol=1e-9, atol=1e-14)
        # global theta index of block b, cell i1: logic_theta * n + i1
        ga = m.logic[a[0], 1] * n + (a[1] // n) % n
        gb = m.logic[b[0], 1] * n + (b[1] // n) % n
        assert np.array_equal(ga, n_th - 1 - gb)


def test_point_source_shell_lengths():
    """Deterministic pin of the tally: per radial shell, sum(J V) over the
    shell's cells equals the packets' exact in-shell ray lengths / 1.001 up to
    the overshoot bookkeeping at shell faces (<= 0.1 % of one step each)."""
    m = _mesh(refine=True)
    src = (0.25, 0.15, 0.1)
    run = RunSpec(bands=[BandSpec(n_packets=3000, kappa_scale=0.0),
                         BandSpec(n_packets=1000, kappa_scale=0.0)], seed=7, step_max=100000)
    out = oracle.sph_forward(m, src, run, weight=2.0, end_state=True, n_threads=4)
    c = out["counters"]
    assert c[_abi.KP_C_PACKETS] == 4000 and c[_abi.KP_C_ESCAPED] == 4000
    J = out["J"].reshape(m.n_cells, 2)
    V = m.cell_volumes()
    # shells: the root-level r faces (every cell lies inside one of them)
    r_f = np.unique(m.xf[m.level == 0][:, 0, :].ravel())
    centers = m.cell_centers()
    shell = np.searchsorted(r_f, centers[:, 0]) - 1
    x0 = np.array(src)
    d = out["end_state"][:, _abi.KP_ES_D:_abi.KP_ES_D + 3]
    # exact length inside [r_k, r_{k+1}] before leaving the domain
    tex = np.zeros((len(r_f) - 1, 2))
    bands = np.r_[np.zeros(3000, int), np.ones(1000, int)]
    for p in range(len(d)):
        ch = _chord(x0, d[p], 0.1, 1.0)
        b = x0 @ d[p]
        # arc-length of the segment [0, ch] inside each shell
        ts = [0.0, ch]
        for rf in r_f:
            det = b * b - (x0 @ x0 - rf * rf)
            if det > 0:
                for t in (-b - np.sqrt(det), -b + np.sqrt(det)):
                    if 0 < t < ch:
                        ts.append(t)
        ts = np.sort(ts)
        mids = 0.5 * (ts[1:] + ts[:-1])
        rm = np.linalg.norm(x0[None, :] + mids[:, None] * d[p][None, :], axis=1)
        k = np.clip(np.searchsorted(r_f, rm) - 1, 0, len(r_f) - 2)
        np.add.at(tex[:, bands[p]], k, np.diff(ts))
    for bnd in range(2):
        got = np.bincount(shell, weights=J[:, bnd] * V, minlength=len(r_f) - 1)
        # weight 2 per packet; J V = w * l
        want = 2.0 * tex[:, bnd] / 1.001
        assert np.allclose(got, want, rtol=2e-3, atol=1e-6), np.max(np.abs(got / want - 1))
        assert got.sum() == pytest.approx(want.sum(), rel=2e-4)
