"""Host block mesh: the reference mesh_t/tree_t/meshgen subset the photon
module consumes (mesh/tree/tree.cpp, mesh/meshgen/meshgen.cpp,
modules/particle/types/proxy.h:22-62), through the C ABI."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi


class Mesh:
    """Block-structured mesh of n_cell^3-cell blocks with 2:1-balanced refinement.

    Mirrors the reference's ``[mesh]`` keys: ``n_cell_global`` is
    ``root_blocks * n_cell``, ``n_cell_block`` is ``n_cell``; ``[boundary]
    kinds`` containing "per" makes an axis periodic; ``[refine_region*]``
    boxes go through :meth:`refine_region`.
    """

    def __init__(self, root_blocks=(4, 4, 4), n_cell=16, x_min=(0.0, 0.0, 0.0),
                 x_max=(1.0, 1.0, 1.0), periodic=(0, 0, 0), coord="cartesian",
                 r_ratio=1.0):
        L = _abi.lib()
        self._h = C.c_void_p()
        rb = (C.c_int32 * 3)(*root_blocks)
        lo = (C.c_double * 3)(*x_min)
        hi = (C.c_double * 3)(*x_max)
        per = (C.c_int32 * 3)(*[int(p) for p in periodic])
        _abi.check(L.kp_mesh_create(rb, int(n_cell), lo, hi, per, C.byref(self._h)))
        self.root_blocks = tuple(int(r) for r in root_blocks)
        self.n_cell = int(n_cell)
        self.x_min = np.array(x_min, dtype=np.float64)
        self.x_max = np.array(x_max, dtype=np.float64)
        self.periodic = tuple(int(p) for p in periodic)
        self.desc = None
        if coord not in ("cartesian", "spherical"):
            raise ValueError("coord must be 'cartesian' or 'spherical'")
        self.coord = coord
        if coord == "spherical":
            _abi.check(L.kp_mesh_set_coord(self._h, _abi.KP_COORD_SPHERICAL, float(r_ratio)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _abi.lib().kp_mesh_destroy(h)
            except (AttributeError, TypeError):
                pass  # interpreter shutdown: module globals are already gone
            self._h = None

    @property
    def handle(self):
        return self._h

    def refine(self, level: int, idx) -> None:
        """tree_t::insert of logical regions ``idx`` ([n,3]) at ``level``."""
        a = np.ascontiguousarray(np.asarray(idx, dtype=np.int32).reshape(-1, 3))
        _abi.check(_abi.lib().kp_mesh_refine(
            self._h, int(level), a.ctypes.data_as(C.POINTER(C.c_int32)), a.shape[0]))

    def refine_region(self, level: int, x_lo, x_hi) -> None:
        lo = (C.c_double * 3)(*x_lo)
        hi = (C.c_double * 3)(*x_hi)
        _abi.check(_abi.lib().kp_mesh_refine_region(self._h, int(level), lo, hi))

    def finalize(self) -> "Mesh":
        L = _abi.lib()
        _abi.check(L.kp_mesh_finalize(self._h))
        d = _abi.MeshDesc()
        _abi.check(L.kp_mesh_describe(self._h, C.byref(d)))
        self.desc = d
        nb = d.n_blocks
        self.n_blocks = nb

        def view(ptr, n, dt):
            return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dt, copy=True)

        self.xf0 = view(d.xf0, 3 * nb, np.float64).reshape(nb, 3)
        self.dx0 = view(d.dx0, 3 * nb, np.float64).reshape(nb, 3)
        self.vol0 = view(d.vol0, nb, np.float64)
        self.level = view(d.level, nb, np.int32)
        self.logic = view(d.logic, 3 * nb, np.int32).reshape(nb, 3)
        self.nb_id = view(d.nb_id, 24 * nb, np.int32).reshape(nb, 24)
        self.nb_shift = view(d.nb_shift, 24 * nb, np.uint8).reshape(nb, 24)
        self.d_lvl = view(d.d_lvl, 6 * nb, np.int8).reshape(nb, 6)
        self.phybnd = view(d.phybnd, nb, np.uint8)
        m = self.n_cell + 1
        self.xf = view(d.xf, 3 * m * nb, np.float64).reshape(nb, 3, m)
        if self.coord == "spherical":
            self.sph = view(d.sph, _abi.sph_tab_len(self.n_cell) * nb, np.float64).reshape(nb, -1)
        return self

    @property
    def n3(self) -> int:
        return self.n_cell ** 3

    @property
    def n_cells(self) -> int:
        return self.n_blocks * self.n3

    def locate(self, x):
        ib = C.c_int32()
        i = (C.c_int32 * 3)()
        _abi.check(_abi.lib().kp_mesh_locate(self._h, (C.c_double * 3)(*x), C.byref(ib), i))
        return ib.value, tuple(i)

    def cell_centers(self) -> np.ndarray:
        """[n_cells, 3] cell-centre coordinates in block-table cell order
        (mesh coordinates: face midpoints per axis, i.e. (r, theta, phi) on a
        spherical mesh)."""
        n = self.n_cell
        if self.coord == "spherical":
            mid = 0.5 * (self.xf[:, :, :-1] + self.xf[:, :, 1:])  # [nb, 3, n]
            c0 = np.broadcast_to(mid[:, 0][:, None, None, :], (self.n_blocks, n, n, n))
            c1 = np.broadcast_to(mid[:, 1][:, None, :, None], (self.n_blocks, n, n, n))
            c2 = np.broadcast_to(mid[:, 2][:, :, None, None], (self.n_blocks, n, n, n))
            return np.stack([c0, c1, c2], axis=-1).reshape(-1, 3)
        i = np.arange(n, dtype=np.float64) + 0.5
        # cell (i0,i1,i2) -> linear i0 + n*(i1 + n*i2)
        g2, g1, g0 = np.meshgrid(i, i, i, indexing="ij")
        loc = np.stack([g0.ravel(), g1.ravel(), g2.ravel()], axis=1)  # [n3, 3]
        return (self.xf0[:, None, :] + loc[None, :, :] * self.dx0[:, None, :]).reshape(-1, 3)

    def cell_volumes(self) -> np.ndarray:
        if self.coord == "spherical":
            n, m = self.n_cell, self.n_cell + 1
            fr = self.sph[:, 4 * m:4 * m + n]
            ft = self.sph[:, 4 * m + n:4 * m + 2 * n]
            fp = self.sph[:, 4 * m + 2 * n:4 * m + 3 * n]
            # cell (i0,i1,i2) -> i0 + n*(i1 + n*i2); volume (f_r*f_th)*f_phi
            v = (fr[:, None, None, :] * ft[:, None, :, None]) * fp[:, :, None, None]
            return v.reshape(-1)
        return np.repeat(self.vol0, self.n3)
