"""Host mesh layer: block tree, ids, geometry and the face-neighbour table.

The table (proxy.h:22-62 semantics) is checked against brute-force point
location of probes just outside every sub-face of every block, so a wrong
neighbour, level difference or coarse-face half fails independently of the
tree code that built it.
"""
import numpy as np
import pytest

from kratos_polrad_amd import KPError
from kratos_polrad_amd.mesh import Mesh


def _random_mesh(seed, root=(3, 2, 2), n_cell=4, periodic=(0, 0, 0), levels=3, n_ref=6):
    rng = np.random.default_rng(seed)
    m = Mesh(root_blocks=root, n_cell=n_cell, x_max=(1.5, 1.0, 1.0), periodic=periodic)
    for lvl in range(1, levels):
        nmax = np.array(root) << lvl
        idx = np.stack([rng.integers(0, nmax[a], n_ref) for a in range(3)], axis=1)
        m.refine(lvl, idx)
    return m.finalize()


def _leaf_at(m, x):
    hi = m.xf0 + m.n_cell * m.dx0
    inside = np.all((x >= m.xf0 - 1e-12) & (x < hi - 1e-12), axis=1)
    b = np.nonzero(inside)[0]
    return int(b[0]) if len(b) == 1 else (-1 if len(b) == 0 else -2)


def _wrap(m, x):
    x = x.copy()
    for a in range(3):
        if m.periodic[a]:
            L = m.x_max[a] - m.x_min[a]
            x[a] = m.x_min[a] + (x[a] - m.x_min[a]) % L
    return x


@pytest.mark.parametrize("seed,periodic", [(1, (0, 0, 0)), (2, (0, 0, 0)), (3, (1, 0, 1)),
                                           (4, (1, 1, 1))])
def test_face_table_matches_geometry(seed, periodic):
    m = _random_mesh(seed, periodic=periodic)
    assert len(set(m.level.tolist())) >= 2
    ext = m.n_cell * m.dx0
    for b in range(m.n_blocks):
        lo, size = m.xf0[b], ext[b]
        for a in range(3):
            ap = [(a + 1) % 3, (a + 2) % 3]
            for side in range(2):
                phy = bool(m.phybnd[b] & (1 << (2 * a + side)))
                dl = int(m.d_lvl[b, 2 * a + side])
                for j0 in range(2):
                    for j1 in range(2):
                        slot = ((a * 2 + side) * 2 + j0) * 2 + j1
                        nb = m.nb_id[b, slot]
                        # probe: just outside the face, at the centre of the
                        # quarter (j0, j1) of the face
                        p = lo.copy()
                        p[a] = lo[a] + (size[a] * (1 + 1e-6) if side else -size[a] * 1e-6)
                        p[ap[0]] = lo[ap[0]] + size[ap[0]] * (0.25 + 0.5 * j0)
                        p[ap[1]] = lo[ap[1]] + size[ap[1]] * (0.25 + 0.5 * j1)
                        outside = p[a] < m.x_min[a] or p[a] >= m.x_max[a]
                        if outside and not m.periodic[a]:
                            assert phy and nb == -1
                            continue
                        assert not phy
                        want = _leaf_at(m, _wrap(m, p))
                        assert want >= 0
                        wl = int(m.level[want]) - int(m.level[b])
                        assert abs(wl) <= 1  # 2:1 balance across faces
                        assert dl == wl
                        if dl > 0 or (j0 == 0 and j1 == 0):
                            assert nb == want, (b, a, side, j0, j1, nb, want)
                        if dl < 0 and j0 == 0 and j1 == 0:
                            # shift bits: which half of the coarse face we touch
                            c = want
                            for j, ax in enumerate(ap):
                                half = (lo[ax] - m.xf0[c, ax]) / ext[c, ax]
                                want_bit = 1 if half >= 0.5 - 1e-9 else 0
                                assert ((m.nb_shift[b, slot] >> j) & 1) == want_bit


def test_two_to_one_balance_everywhere():
    m = _random_mesh(7, levels=4, n_ref=4)
    ext = m.n_cell * m.dx0
    for b in range(m.n_blocks):
        for off in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1])).T.reshape(-1, 3):
            if not off.any():
                continue
            p = m.xf0[b] + ext[b] * (0.5 + 0.5 * off * (1 + 1e-6))
            if np.any(p < m.x_min) or np.any(p >= m.x_max):
                continue
            nb = _leaf_at(m, p)
            assert abs(int(m.level[nb]) - int(m.level[b])) <= 1


def test_ids_follow_lexicographic_order():
    """dist_default (meshgen.cpp:168-188): ids in (level, x0, x1, x2) order."""
    m = _random_mesh(5)
    keys = [(int(m.level[b]), *m.logic[b].tolist()) for b in range(m.n_blocks)]
    assert keys == sorted(keys)


def test_geometry_and_volume():
    m = Mesh(root_blocks=(2, 2, 2), n_cell=8, x_min=(-1, 0, 0), x_max=(1, 1, 2))
    m.refine(1, [[1, 1, 1]])
    m.finalize()
    # leaf volume covers the domain exactly
    assert abs(np.sum(m.vol0 * m.n3) - 4.0) < 1e-12
    for b in range(m.n_blocks):
        lvl = m.level[b]
        want = np.array([2.0, 1.0, 2.0]) / (np.array([2, 2, 2]) << lvl) / 8
        assert np.allclose(m.dx0[b], want)
        assert np.allclose(m.xf0[b], np.array([-1, 0, 0]) + m.logic[b] * want * 8)


def test_locate_half_open_cells():
    m = Mesh(root_blocks=(4, 4, 4), n_cell=16).finalize()
    ib, i = m.locate((0.5, 0.5, 0.5))
    assert tuple(m.logic[ib]) == (2, 2, 2) and i == (0, 0, 0)
    ib, i = m.locate((0.499999, 0.25, 0.75))
    assert tuple(m.logic[ib]) == (1, 1, 3) and i == (15, 0, 0)
    ib, _ = m.locate((1.0, 0.5, 0.5))
    assert ib == -1


def test_errors_are_reported():
    with pytest.raises(KPError, match="even"):
        Mesh(root_blocks=(1, 1, 1), n_cell=5)
    m = Mesh(root_blocks=(2, 2, 2), n_cell=4)
    with pytest.raises(KPError, match="outside"):
        m.refine(1, [[4, 0, 0]])
    with pytest.raises(KPError, match="SMR"):
        m.refine_region(1, (-1, 0, 0), (0.5, 0.5, 0.5))
    m.finalize()
    with pytest.raises(KPError, match="finalized"):
        m.refine(1, [[0, 0, 0]])
