"""Transport engine through the C ABI (include/kpolrad.h).

Device outputs are torch tensors on the engine's GPU; torch provides device
memory and the stream, the transport itself runs in libkpolrad.so."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _abi


@dataclass
class BandSpec:
    """One wavelength / source set (a component of the rad_t::flx tally)."""
    n_packets: int
    kappa_scale: float = 1.0
    albedo: float = 1.0
    g: float = 0.0
    cell_albedo: bool = False
    cell_g: bool = False


@dataclass
class RunSpec:
    bands: List[BandSpec]
    seed: int = 42
    scatter_model: int = _abi.KP_THOMSON
    p_max: float = 1.0
    precision: int = _abi.KP_F64
    packet_offset: int = 0
    step_max: int = 1 << 30
    incl_deg: Sequence[float] = ()
    rr_weight: float = 1e-2
    rr_survive: float = 0.1
    image_n: int = 128
    image_half: float = 0.5

    def to_c(self):
        """ctypes kp_run plus the arrays it points to (keep them alive)."""
        bands = (_abi.Band * len(self.bands))()
        for k, b in enumerate(self.bands):
            bands[k].kappa_scale = float(b.kappa_scale)
            bands[k].albedo = float(b.albedo)
            bands[k].g = float(b.g)
            bands[k].n_packets = int(b.n_packets)
            bands[k].flags = ((_abi.KP_BAND_CELL_ALBEDO if b.cell_albedo else 0)
                              | (_abi.KP_BAND_CELL_G if b.cell_g else 0))
        incl = (C.c_double * max(1, len(self.incl_deg)))(*self.incl_deg)
        r = _abi.Run()
        r.seed = int(self.seed) & 0xFFFFFFFF
        r.scatter_model = int(self.scatter_model)
        r.p_max = float(self.p_max)
        r.n_bands = len(self.bands)
        r.precision = int(self.precision)
        r.bands = C.cast(bands, C.POINTER(_abi.Band))
        r.packet_offset = int(self.packet_offset)
        r.step_max = int(self.step_max)
        r.n_incl = len(self.incl_deg)
        r.incl_deg = C.cast(incl, C.POINTER(C.c_double))
        r.rr_weight = float(self.rr_weight)
        r.rr_survive = float(self.rr_survive)
        r.image_n = int(self.image_n)
        r.image_half = float(self.image_half)
        return r, (bands, incl)

    @property
    def total_packets(self) -> int:
        return sum(int(b.n_packets) for b in self.bands)


def source_c(x, weight=1.0):
    s = _abi.Source()
    for a in range(3):
        s.x[a] = float(x[a])
    s.weight = float(weight)
    return s


class Engine:
    def __init__(self, mesh, device: int = 0):
        import torch  # device memory and streams

        if mesh.desc is None:
            raise ValueError("mesh must be finalized")
        self.mesh = mesh
        self.device = int(device)
        self._torch = torch
        self._h = C.c_void_p()
        _abi.check(_abi.lib().kp_engine_create(C.byref(mesh.desc), self.device,
                                                 C.byref(self._h)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _abi.lib().kp_engine_destroy(h)
            except (AttributeError, TypeError):
                pass  # interpreter shutdown: module globals are already gone
            self._h = None

    @property
    def handle(self):
        return self._h

    def set_schedule(self, kind=_abi.KP_SCHED_PERSISTENT, block_threads=256, blocks_per_cu=4,
                     chunk=64, j_replicas=0, img_replicas=0, event_batch=8,
                     waves_per_simd=0, coherent=0, peel_queue=1, vertex_slots=0,
                     adjoint_chunk=0, queue_from_band=0, queue_passes=0,
                     queue_chunk=0, queue_yield=0) -> None:
        s = _abi.Schedule(int(kind), int(block_threads), int(blocks_per_cu), int(chunk),
                          int(j_replicas), int(img_replicas), int(event_batch),
                          int(waves_per_simd), int(coherent), int(peel_queue),
                          int(vertex_slots), int(adjoint_chunk), int(queue_from_band),
                          int(queue_passes), int(queue_chunk), int(queue_yield))
        _abi.check(_abi.lib().kp_set_schedule(self._h, C.byref(s)))

    def get_schedule(self) -> dict:
        s = _abi.Schedule()
        _abi.check(_abi.lib().kp_get_schedule(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _abi.Schedule._fields_}

    def autotune(self, source_x, run: "RunSpec", sample_packets=2_000_000, force=False,
                 weight=1.0, stream=None) -> dict:
        """Time candidate schedules on a packet sample and install the fastest."""
        t = self._torch
        r, keep = run.to_c()
        best = _abi.Schedule()
        src = source_c(source_x, weight)
        if stream is None:
            stream = t.cuda.current_stream(self.device)
        _abi.check(_abi.lib().kp_autotune(self._h, C.byref(src), C.byref(r), int(sample_packets),
                                           int(bool(force)), C.byref(best),
                                           C.c_void_p(stream.cuda_stream)))
        del keep
        return {k: getattr(best, k) for k, _ in _abi.Schedule._fields_}

    def graph(self, source_x, run: "RunSpec", outputs: dict, n_chunks=1, weight=1.0):
        """Capture a forward run (n_chunks chunks) into a replayable hipGraph."""
        return Graph(self, source_x, run, outputs, n_chunks, weight)

    def set_media(self, kappa_rho, albedo, g=None) -> None:
        n = self.mesh.n_cells
        k = np.ascontiguousarray(np.asarray(kappa_rho, dtype=np.float32).reshape(n))
        a = np.ascontiguousarray(np.asarray(albedo, dtype=np.float32).reshape(n))
        gp = None
        if g is not None:
            gg = np.ascontiguousarray(np.asarray(g, dtype=np.float32).reshape(n))
            gp = gg.ctypes.data_as(C.c_void_p)
        _abi.check(_abi.lib().kp_set_media(self._h, k.ctypes.data_as(C.c_void_p),
                                            a.ctypes.data_as(C.c_void_p), gp))

    def alloc_outputs(self, run: RunSpec, J=True, image=True, counters=True,
                      end_state=False):
        t = self._torch
        dev = t.device("cuda", self.device)
        out = {}
        nb = len(run.bands)
        if J:
            out["J"] = t.zeros(self.mesh.n_cells * nb, dtype=t.float64, device=dev)
        if image and len(run.incl_deg):
            out["image"] = t.zeros((nb, len(run.incl_deg), run.image_n, run.image_n, 3),
                                   dtype=t.float64, device=dev)
        if counters:
            out["counters"] = t.zeros(_abi.KP_NCOUNTERS, dtype=t.int64, device=dev)
        if end_state:
            out["end_state"] = t.zeros((run.total_packets, _abi.KP_ES_STRIDE),
                                       dtype=t.float64, device=dev)
        return out

    def forward(self, source_x, run: RunSpec, outputs: dict, weight=1.0, stream=None):
        """Accumulate one launch into ``outputs`` (tensors from alloc_outputs)."""
        t = self._torch
        r, keep = run.to_c()
        tal = _abi.Tallies()
        for key in ("J", "image", "counters", "end_state"):
            v = outputs.get(key)
            setattr(tal, key, v.data_ptr() if v is not None else None)
        src = source_c(source_x, weight)
        if stream is None:
            stream = t.cuda.current_stream(self.device)
        _abi.check(_abi.lib().kp_forward(self._h, C.byref(src), C.byref(r), C.byref(tal),
                                          C.c_void_p(stream.cuda_stream)))
        del keep
        return outputs

    def adjoint(self, source_x, run: RunSpec, dL_dimage, weight=1.0, grads=None, stream=None):
        """Path-replay gradients of sum(dL_dimage * image) (FP64).

        dL_dimage: device tensor in the image layout.  Returns a dict of device
        tensors d_logrho, d_albedo, d_g (per cell) and jlin (scalar), added
        into ``grads`` if given."""
        t = self._torch
        dev = t.device("cuda", self.device)
        if grads is None:
            grads = {k: t.zeros(self.mesh.n_cells, dtype=t.float64, device=dev)
                     for k in ("d_logrho", "d_albedo", "d_g")}
            grads["jlin"] = t.zeros(1, dtype=t.float64, device=dev)
        r, keep = run.to_c()
        g = _abi.Grads()
        for key in ("d_logrho", "d_albedo", "d_g", "jlin"):
            v = grads.get(key)
            setattr(g, key, v.data_ptr() if v is not None else None)
        w = dL_dimage.contiguous()
        src = source_c(source_x, weight)
        if stream is None:
            stream = t.cuda.current_stream(self.device)
        _abi.check(_abi.lib().kp_adjoint(self._h, C.byref(src), C.byref(r),
                                          C.c_void_p(w.data_ptr()), C.byref(g),
                                          C.c_void_p(stream.cuda_stream)))
        del keep
        return grads


class Graph:
    """A captured forward run; launch() replays it (tallies accumulate)."""

    def __init__(self, engine, source_x, run, outputs, n_chunks=1, weight=1.0):
        self.engine = engine
        self.outputs = outputs
        r, keep = run.to_c()
        tal = _abi.Tallies()
        for key in ("J", "image", "counters", "end_state"):
            v = outputs.get(key)
            setattr(tal, key, v.data_ptr() if v is not None else None)
        src = source_c(source_x, weight)
        self._h = C.c_void_p()
        _abi.check(_abi.lib().kp_graph_create(engine.handle, C.byref(src), C.byref(r),
                                               C.byref(tal), int(n_chunks), C.byref(self._h)))
        del keep

    def launch(self, stream=None):
        t = self.engine._torch
        if stream is None:
            stream = t.cuda.current_stream(self.engine.device)
        _abi.check(_abi.lib().kp_graph_launch(self._h, C.c_void_p(stream.cuda_stream)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _abi.lib().kp_graph_destroy(h)
            except (AttributeError, TypeError):
                pass  # interpreter shutdown: module globals are already gone
            self._h = None
