"""FP64 CPU oracle of the photon-packet transport path -- TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py load
this package, as the checker or the timed CPU baseline.  The product
(kratos_polrad_amd) never imports it.

The C sources restate the reference march of
src/modules/particle/radiation/photon.h:20-125, types/proxy.h:64-121,
types/particle_base.h:37-144 and mesh/geometry/geometry.h:68-83, plus the
physics the reference has no code for: Philox sampling, polarized
scattering, absorption weighting, peel-off images (see kp_oracle.h for the
list and the documented deviations), and the spherical march of
radiation/photon_sph.h:10-266 (kp_sph_oracle.c).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# KP_ORACLE_PLANT=k (k > 0) builds and loads the oracle with the k-th planted
# mistake (kp_oracle.c, KP_PLANT) from KP_ORACLE_OUT, never over the real one
PLANT = int(os.environ.get("KP_ORACLE_PLANT", "0") or 0)
LIB_PATH = os.path.join(HERE, "liboracle.so")
if PLANT:
    LIB_PATH = os.environ.get("KP_ORACLE_OUT", "")
    if not LIB_PATH or os.path.dirname(os.path.abspath(LIB_PATH)) == HERE:
        raise RuntimeError("KP_ORACLE_PLANT needs KP_ORACLE_OUT outside oracle/")

_lib = None


def build() -> str:
    cmd = ["make", "-s", "-C", HERE]
    if PLANT:
        cmd += [f"PLANT={PLANT}", f"OUT={os.path.abspath(LIB_PATH)}"]
    subprocess.run(cmd, check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        from kratos_polrad_amd import _abi

        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.or_forward.restype = C.c_int
        L.or_forward.argtypes = [C.POINTER(_abi.MeshDesc), P, P, P, C.POINTER(_abi.Source),
                                 C.POINTER(_abi.Run), P, P, P, P, C.c_int]
        L.or_trace.restype = C.c_int
        L.or_trace.argtypes = [C.POINTER(_abi.MeshDesc), C.POINTER(C.c_double),
                               C.POINTER(C.c_double), C.c_int, P, P, P,
                               C.POINTER(C.c_double), C.POINTER(C.c_int)]
        L.or_locate.restype = C.c_int
        L.or_locate.argtypes = [C.POINTER(_abi.MeshDesc), C.POINTER(C.c_double),
                                C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.or_philox.restype = None
        L.or_philox.argtypes = [C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32,
                                C.POINTER(C.c_uint32)]
        for name in ("or_log", "or_exp", "or_thomson_mu"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [C.c_double]
        L.or_hg_mu.restype = C.c_double
        L.or_hg_mu.argtypes = [C.c_double, C.c_double]
        L.or_sincos2pi.restype = None
        L.or_sincos2pi.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_adjoint.restype = C.c_int
        L.or_adjoint.argtypes = [C.POINTER(_abi.MeshDesc), P, P, P, C.POINTER(_abi.Source),
                                 C.POINTER(_abi.Run), P, P, P, P, C.POINTER(C.c_double), C.c_int]
        D3 = C.POINTER(C.c_double)
        L.or_scatter_stokes.restype = None
        L.or_scatter_stokes.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, D3, D3, D3]
        L.or_peel_stokes.restype = None
        L.or_peel_stokes.argtypes = [C.c_int, C.c_double, C.c_double, D3, D3, D3, D3, D3, D3]
        L.or_pool_keep.restype = C.c_int64
        L.or_pool_keep.argtypes = [P, C.c_int64, C.c_int64, P]
        L.or_bin_by_key.restype = C.c_int64
        L.or_bin_by_key.argtypes = [P, C.c_int64, C.c_int32, P, P, P]
        L.or_gather.restype = None
        L.or_gather.argtypes = [P, C.c_int64, P, C.c_int64, P]
        L.or_sph_forward.restype = C.c_int
        L.or_sph_forward.argtypes = [C.POINTER(_abi.MeshDesc), C.POINTER(_abi.Source),
                                     C.POINTER(_abi.Run), P, P, P, C.c_int]
        L.or_last_transport_seconds.restype = C.c_double
        L.or_last_transport_seconds.argtypes = []
        L.or_sph_trace.restype = C.c_int
        L.or_sph_trace.argtypes = L.or_trace.argtypes
        L.or_sph_locate.restype = C.c_int
        L.or_sph_locate.argtypes = L.or_locate.argtypes
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def forward(mesh, kappa, albedo, g, source_x, run_spec, weight=1.0, J=True, image=True,
            end_state=False, n_threads=None):
    """Run the oracle on the same inputs as Engine.forward; returns numpy outputs."""
    from kratos_polrad_amd import _abi
    from kratos_polrad_amd.engine import source_c

    n = mesh.n_cells
    k = np.ascontiguousarray(np.asarray(kappa, dtype=np.float32).reshape(n))
    a = np.ascontiguousarray(np.asarray(albedo, dtype=np.float32).reshape(n))
    gg = None if g is None else np.ascontiguousarray(np.asarray(g, dtype=np.float32).reshape(n))
    r, keep = run_spec.to_c()
    nb = len(run_spec.bands)
    out = {"counters": np.zeros(_abi.KP_NCOUNTERS, dtype=np.uint64)}
    if J:
        out["J"] = np.zeros(n * nb, dtype=np.float64)
    if image and len(run_spec.incl_deg):
        out["image"] = np.zeros((nb, len(run_spec.incl_deg), run_spec.image_n,
                                 run_spec.image_n, 3), dtype=np.float64)
    if end_state:
        out["end_state"] = np.zeros((run_spec.total_packets, _abi.KP_ES_STRIDE))
    src = source_c(source_x, weight)
    if n_threads is None:
        n_threads = os.cpu_count() or 1
    rc = lib().or_forward(C.byref(mesh.desc), _ptr(k), _ptr(a), _ptr(gg), C.byref(src),
                          C.byref(r), _ptr(out.get("J")), _ptr(out.get("image")),
                          _ptr(out["counters"]), _ptr(out.get("end_state")), int(n_threads))
    del keep
    if rc != 0:
        raise RuntimeError(f"or_forward failed: {rc}")
    out["transport_seconds"] = float(lib().or_last_transport_seconds())
    return out


def trace(mesh, x, d, max_steps=100000):
    """Pure-streaming march of one ray: (ib[], cell[], length[], x_end, status)."""
    ib = np.zeros(max_steps, dtype=np.int32)
    cell = np.zeros(max_steps, dtype=np.int32)
    ln = np.zeros(max_steps, dtype=np.float64)
    xe = (C.c_double * 3)()
    st = C.c_int()
    k = lib().or_trace(C.byref(mesh.desc), (C.c_double * 3)(*x), (C.c_double * 3)(*d),
                       max_steps, _ptr(ib), _ptr(cell), _ptr(ln), xe, C.byref(st))
    return ib[:k], cell[:k], ln[:k], np.array(xe[:]), st.value


def sph_forward(mesh, source_x, run_spec, weight=1.0, J=True, end_state=False,
                n_threads=None):
    """Spherical-mesh streaming run (or_sph_forward) on the same inputs as
    Engine.forward on a spherical mesh; returns numpy outputs."""
    from kratos_polrad_amd import _abi
    from kratos_polrad_amd.engine import source_c

    r, keep = run_spec.to_c()
    nb = len(run_spec.bands)
    out = {"counters": np.zeros(_abi.KP_NCOUNTERS, dtype=np.uint64)}
    if J:
        out["J"] = np.zeros(mesh.n_cells * nb, dtype=np.float64)
    if end_state:
        out["end_state"] = np.zeros((run_spec.total_packets, _abi.KP_ES_STRIDE))
    src = source_c(source_x, weight)
    if n_threads is None:
        n_threads = os.cpu_count() or 1
    rc = lib().or_sph_forward(C.byref(mesh.desc), C.byref(src), C.byref(r), _ptr(out.get("J")),
                              _ptr(out["counters"]), _ptr(out.get("end_state")), int(n_threads))
    del keep
    if rc != 0:
        raise RuntimeError(f"or_sph_forward failed: {rc}")
    return out


def sph_trace(mesh, x, d, max_steps=100000):
    """Spherical march of one ray from Cartesian x along d:
    (ib[], cell[], length[], x_end, status)."""
    ib = np.zeros(max_steps, dtype=np.int32)
    cell = np.zeros(max_steps, dtype=np.int32)
    ln = np.zeros(max_steps, dtype=np.float64)
    xe = (C.c_double * 3)()
    st = C.c_int()
    k = lib().or_sph_trace(C.byref(mesh.desc), (C.c_double * 3)(*x), (C.c_double * 3)(*d),
                           max_steps, _ptr(ib), _ptr(cell), _ptr(ln), xe, C.byref(st))
    return ib[:k], cell[:k], ln[:k], np.array(xe[:]), st.value


def sph_locate(mesh, x):
    """Cell of Cartesian point x on a spherical mesh: (ib, (i0, i1, i2)), ib=-1 outside."""
    ib = C.c_int32()
    i = (C.c_int32 * 3)()
    lib().or_sph_locate(C.byref(mesh.desc), (C.c_double * 3)(*x), C.byref(ib), i)
    return ib.value, tuple(i)


def philox(ctr, k0, k1):
    c = (C.c_uint32 * 4)(*ctr)
    o = (C.c_uint32 * 4)()
    lib().or_philox(c, k0, k1, o)
    return list(o)


def scatter_stokes(model, pmax, mu, phi_u, d, e1, S):
    """One scattering's Stokes/frame algebra: returns (d', e1', S')."""
    dd = (C.c_double * 3)(*d)
    ee = (C.c_double * 3)(*e1)
    ss = (C.c_double * 3)(*S)
    lib().or_scatter_stokes(int(model), float(pmax), float(mu), float(phi_u), dd, ee, ss)
    return np.array(dd[:]), np.array(ee[:]), np.array(ss[:])


def peel_stokes(model, g, pmax, d, e1, S, no, ev):
    """Stokes vector peeled toward no, in the image frame (Q>0 along ev)."""
    out = (C.c_double * 3)()
    lib().or_peel_stokes(int(model), float(g), float(pmax), (C.c_double * 3)(*d),
                         (C.c_double * 3)(*e1), (C.c_double * 3)(*S), (C.c_double * 3)(*no),
                         (C.c_double * 3)(*ev), out)
    return np.array(out[:])


def adjoint(mesh, kappa, albedo, g, source_x, run_spec, dL_dimage, weight=1.0, n_threads=None):
    """FP64 path-replay adjoint of Jlin = sum(dL_dimage * image): returns
    (d_logrho, d_albedo, d_g, Jlin) as numpy arrays / float."""
    from kratos_polrad_amd.engine import source_c

    n = mesh.n_cells
    k = np.ascontiguousarray(np.asarray(kappa, dtype=np.float32).reshape(n))
    a = np.ascontiguousarray(np.asarray(albedo, dtype=np.float32).reshape(n))
    gg = None if g is None else np.ascontiguousarray(np.asarray(g, dtype=np.float32).reshape(n))
    w = np.ascontiguousarray(np.asarray(dL_dimage, dtype=np.float64))
    r, keep = run_spec.to_c()
    dl = np.zeros(n)
    da = np.zeros(n)
    dg = np.zeros(n)
    jl = C.c_double(0.0)
    src = source_c(source_x, weight)
    if n_threads is None:
        n_threads = os.cpu_count() or 1
    rc = lib().or_adjoint(C.byref(mesh.desc), _ptr(k), _ptr(a), _ptr(gg), C.byref(src),
                          C.byref(r), _ptr(w), _ptr(dl), _ptr(da), _ptr(dg), C.byref(jl),
                          int(n_threads))
    del keep
    if rc != 0:
        raise RuntimeError(f"or_adjoint failed: {rc}")
    return dl, da, dg, jl.value


def pool_keep(pool: np.ndarray, keep: np.ndarray):
    """In-place keep_pool compaction of the records (rows) of ``pool``; returns
    (compacted copy, n_keep)."""
    p = np.ascontiguousarray(pool).copy()
    k = np.ascontiguousarray(keep, dtype=np.uint8)
    n = p.shape[0]
    unit = p.strides[0] if n else 0
    nk = lib().or_pool_keep(_ptr(p), n, unit, _ptr(k))
    return p, int(nk)


def bin_by_key(keys: np.ndarray, n_bins: int):
    """Stable per-destination binning: (offsets[n_bins+1], order[n_sent], slot[n])."""
    k = np.ascontiguousarray(keys, dtype=np.int32)
    n = k.shape[0]
    offsets = np.zeros(n_bins + 1, dtype=np.int64)
    order = np.zeros(max(n, 1), dtype=np.int32)
    slot = np.zeros(max(n, 1), dtype=np.int32)
    ns = lib().or_bin_by_key(_ptr(k), n, int(n_bins), _ptr(offsets), _ptr(order), _ptr(slot))
    return offsets, order[:ns].copy(), slot[:n].copy()


def gather(src: np.ndarray, order: np.ndarray):
    s = np.ascontiguousarray(src)
    o = np.ascontiguousarray(order, dtype=np.int32)
    out = np.empty((len(o),) + s.shape[1:], dtype=s.dtype)
    unit = s.strides[0] if len(s) else 0
    lib().or_gather(_ptr(s), unit, _ptr(o), len(o), _ptr(out))
    return out
