"""Batched multi-band launches, hipGraph replay and the schedule autotuner.

Multi-band FP64 runs are checked bit-exactly against the oracle (uneven
packet counts, per-band extinction scale, albedo and g, an empty band).
Graph replays must reproduce the eager launches of the same chunks: FP64
end states bit-identical, counters identical, tallies equal up to the order
of FP64 atomic sums; replays are interleaved with eager work on the same
engine and every replay is checked.
"""
import numpy as np
import pytest

from kratos_polrad_amd import _abi, configs
from kratos_polrad_amd.engine import BandSpec, Engine, RunSpec
from kratos_polrad_amd.mesh import Mesh

pytestmark = pytest.mark.gpu


def _bands_workload(precision=_abi.KP_F64, counts=(1500, 0, 400, 2600)):
    mesh = Mesh(root_blocks=(4, 4, 4), n_cell=8)
    mesh.refine_region(1, (0.3, 0.3, 0.3), (0.7, 0.7, 0.7))
    mesh.refine_region(2, (0.42, 0.45, 0.4), (0.58, 0.55, 0.6))
    mesh.finalize()
    xc = mesh.cell_centers()
    r2 = np.sum((xc - 0.5) ** 2, axis=1)
    kappa = (0.6 / (r2 + 0.01)).astype(np.float32)
    kappa[r2 > 0.2] = 0.0
    albedo = np.full(mesh.n_cells, 0.5, dtype=np.float32)
    g = np.full(mesh.n_cells, 0.6, dtype=np.float32)
    ks = (2.8, 1.0, 0.4, 0.06)
    gs = (0.7, 0.6, 0.45, 0.3)
    bands = [BandSpec(n_packets=n, kappa_scale=k, albedo=0.5, g=gg)
             for n, k, gg in zip(counts, ks, gs)]
    run = RunSpec(bands=bands, seed=99, scatter_model=_abi.KP_HG_RAYLEIGH, p_max=0.5,
                  precision=precision, incl_deg=(90.0, 30.0), image_n=32, image_half=0.6)
    return configs.Workload("bands_test", mesh, kappa, albedo, g, (0.5, 0.52, 0.49), run)


def _host(out):
    import torch

    torch.cuda.synchronize()
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res["counters"] = res["counters"].astype(np.uint64)
    return res


def _zero(out):
    for v in out.values():
        v.zero_()  # a fill kernel, outside any graph


def _assert_same(a, b, exact_es=True):
    assert np.array_equal(a["counters"], b["counters"])
    if exact_es and "end_state" in a:
        np.testing.assert_array_equal(a["end_state"], b["end_state"])
    np.testing.assert_allclose(a["J"], b["J"], rtol=1e-12, atol=1e-12 * np.abs(b["J"]).max())
    np.testing.assert_allclose(a["image"], b["image"], rtol=1e-11,
                               atol=1e-12 * np.abs(b["image"]).max())


def test_multiband_f64_bit_exact(oracle_lib):
    w = _bands_workload()
    eng = Engine(w.mesh)
    eng.set_media(w.kappa, w.albedo, w.g)
    out = eng.alloc_outputs(w.run, end_state=True)
    eng.forward(w.source, w.run, out)
    got = _host(out)
    ref = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, w.run, end_state=True)
    _assert_same(got, ref)
    # bands are independent Philox streams keyed by band index: a run in
    # which only band 2 has packets reproduces band 2 of the full run
    solo = RunSpec(**{**w.run.__dict__, "bands": [
        BandSpec(**{**b.__dict__, "n_packets": (b.n_packets if k == 2 else 0)})
        for k, b in enumerate(w.run.bands)]})
    o1 = eng.alloc_outputs(solo, end_state=True)
    eng.forward(w.source, solo, o1)
    h1 = _host(o1)
    s0 = sum(b.n_packets for b in w.run.bands[:2])
    np.testing.assert_array_equal(h1["end_state"], got["end_state"][s0:s0 + 400])
    J4 = got["J"].reshape(-1, 4)
    np.testing.assert_allclose(h1["J"].reshape(-1, 4)[:, 2], J4[:, 2], rtol=1e-12,
                               atol=1e-12 * np.abs(J4[:, 2]).max())


@pytest.mark.parametrize("precision", [_abi.KP_F64, _abi.KP_F32], ids=["f64", "f32"])
def test_graph_replays_match_eager(precision):
    import torch

    w = _bands_workload(precision)
    eng = Engine(w.mesh)
    eng.set_media(w.kappa, w.albedo, w.g)
    n_chunks = 3
    stride = max(b.n_packets for b in w.run.bands)
    # eager reference: the same chunks as separate launches
    ref_out = eng.alloc_outputs(w.run)
    for k in range(n_chunks):
        r = RunSpec(**{**w.run.__dict__, "packet_offset": w.run.packet_offset + k * stride})
        eng.forward(w.source, r, ref_out)
    ref = _host(ref_out)
    out = eng.alloc_outputs(w.run)
    g = eng.graph(w.source, w.run, out, n_chunks=n_chunks)
    other = RunSpec(**{**w.run.__dict__, "bands": [BandSpec(n_packets=777, albedo=0.9)],
                       "packet_offset": 1 << 33})
    scratch = eng.alloc_outputs(other)
    for rep in range(3):
        g.launch()
        got = _host(out)
        if precision == _abi.KP_F64:
            _assert_same(got, ref, exact_es=False)
        else:
            assert np.array_equal(got["counters"], ref["counters"])
            np.testing.assert_allclose(got["J"], ref["J"], rtol=1e-6,
                                       atol=1e-9 * np.abs(ref["J"]).max())
            np.testing.assert_allclose(got["image"], ref["image"], rtol=1e-6,
                                       atol=1e-9 * np.abs(ref["image"]).max())
        _zero(out)
        # eager work on the same engine between replays (other band table,
        # other replica shapes)
        eng.forward(w.source, other, scratch)
        torch.cuda.synchronize()
    del g


def test_graph_end_state_single_chunk(oracle_lib):
    w = _bands_workload()
    eng = Engine(w.mesh)
    eng.set_media(w.kappa, w.albedo, w.g)
    out = eng.alloc_outputs(w.run, end_state=True)
    g = eng.graph(w.source, w.run, out, n_chunks=1)
    g.launch()
    got = _host(out)
    ref = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, w.run, end_state=True)
    _assert_same(got, ref)
    with pytest.raises(_abi.KPError):
        eng.graph(w.source, w.run, out, n_chunks=2)


def test_graph_sees_new_medium():
    """A graph reads the engine's medium, which kp_set_media rewrites in place."""
    w = _bands_workload()
    eng = Engine(w.mesh)
    eng.set_media(w.kappa, w.albedo, w.g)
    out = eng.alloc_outputs(w.run)
    g = eng.graph(w.source, w.run, out)
    k2 = w.kappa * 1.5
    eng.set_media(k2, w.albedo, w.g)
    g.launch()
    got = _host(out)
    ref_out = eng.alloc_outputs(w.run)
    eng.forward(w.source, w.run, ref_out)
    _assert_same(got, _host(ref_out), exact_es=False)


def test_autotune_installs_valid_schedule(oracle_lib):
    w = _bands_workload(counts=(3000, 0, 800, 5000))
    eng = Engine(w.mesh)
    eng.set_media(w.kappa, w.albedo, w.g)
    best = eng.autotune(w.source, w.run, sample_packets=4000)
    assert eng.get_schedule() == best
    assert best["kind"] == _abi.KP_SCHED_PERSISTENT
    assert best["block_threads"] in (128, 256)
    # cached: a second call returns the same schedule without re-timing
    assert eng.autotune(w.source, w.run, sample_packets=4000) == best
    # FP64 results do not depend on the schedule
    out = eng.alloc_outputs(w.run, end_state=True)
    eng.forward(w.source, w.run, out)
    ref = oracle_lib.forward(w.mesh, w.kappa, w.albedo, w.g, w.source, w.run, end_state=True)
    _assert_same(_host(out), ref)
