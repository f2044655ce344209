"""The photon module through its C ABI (include/kp_module.h).

PhotonModule mirrors the reference's particle::radiation::base_t life cycle
(read -> init -> step* -> save) for Python callers; read_container parses the
binary_io files that kp_module_save writes (format of the reference's
io/binary/binary_io.h:6-32 with the encodings of
utilities/data_transfer/serialize.h)."""
from __future__ import annotations

import ctypes as C
import struct

import numpy as np

from . import _abi


class PhotonModule:
    def __init__(self, ini_text: str):
        self._h = C.c_void_p()
        _abi.check(_abi.lib().kp_module_create(ini_text.encode(), C.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _abi.lib().kp_module_destroy(h)
            except (AttributeError, TypeError):
                pass  # interpreter shutdown: module globals are already gone
            self._h = None

    def init(self, device: int = 0) -> "PhotonModule":
        """Builds the mesh; with device >= 0 also the engine (GPU)."""
        _abi.check(_abi.lib().kp_module_init(self._h, int(device)))
        return self

    def info(self) -> dict:
        i = _abi.ModuleInfo()
        _abi.check(_abi.lib().kp_module_get_info(self._h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in _abi.ModuleInfo._fields_}

    def set_media(self, kappa_rho, albedo, g=None) -> None:
        n = self.info()["n_cells"]
        k = np.ascontiguousarray(np.asarray(kappa_rho, dtype=np.float32).reshape(n))
        a = np.ascontiguousarray(np.asarray(albedo, dtype=np.float32).reshape(n))
        gg = None if g is None else np.ascontiguousarray(np.asarray(g, np.float32).reshape(n))
        _abi.check(_abi.lib().kp_module_set_media(
            self._h, k.ctypes.data_as(C.c_void_p), a.ctypes.data_as(C.c_void_p),
            None if gg is None else gg.ctypes.data_as(C.c_void_p)))

    def step(self, stream=None) -> None:
        s = None if stream is None else C.c_void_p(stream.cuda_stream)
        _abi.check(_abi.lib().kp_module_step(self._h, s))

    def save(self, path: str) -> None:
        _abi.check(_abi.lib().kp_module_save(self._h, str(path).encode()))

    def tallies(self) -> dict:
        i = self.info()
        J = np.zeros((i["n_cells"], i["n_bands"]))
        img = np.zeros((i["n_bands"], i["n_incl"], i["image_n"], i["image_n"], 3))
        cnt = np.zeros(_abi.KP_NCOUNTERS, dtype=np.uint64)
        _abi.check(_abi.lib().kp_module_tallies(
            self._h, J.ctypes.data_as(C.c_void_p),
            img.ctypes.data_as(C.c_void_p) if img.size else None,
            cnt.ctypes.data_as(C.c_void_p)))
        return {"J": J, "image": img, "counters": cnt}


def read_container(path: str) -> dict:
    """tag -> (bytes, u_size) of a binary_io container file."""
    raw = open(path, "rb").read()
    lead = raw[0]
    if not lead & 1 or (lead & ~1) != 8:
        raise ValueError("only little-endian files with 8-byte size_t are read here")
    (hsize,) = struct.unpack_from("<Q", raw, 1)
    pos = 9
    (n,) = struct.unpack_from("<Q", raw, pos)
    pos += 8
    info = {}
    for _ in range(n):
        (ln,) = struct.unpack_from("<Q", raw, pos)
        pos += 8
        tag = raw[pos:pos + ln].decode()
        pos += ln
        size, u_size, off = struct.unpack_from("<QbQ", raw, pos)
        pos += 17
        info[tag] = (size, u_size, off)
    if pos != hsize:
        raise ValueError("inconsistent header size")
    return {t: (raw[hsize + off: hsize + off + size], u) for t, (size, u, off) in info.items()}


def read_string_map(blob: bytes) -> dict:
    """serialize::read of std::map<std::string, std::string>."""
    (n,) = struct.unpack_from("<Q", blob, 0)
    pos, out = 8, {}
    for _ in range(n):
        vals = []
        for _k in range(2):
            (ln,) = struct.unpack_from("<Q", blob, pos)
            pos += 8
            vals.append(blob[pos:pos + ln].decode())
            pos += ln
        out[vals[0]] = vals[1]
    return out
